/*
 * bm.h -- C ABI of the MI355X (gfx950) exact squared EDT and local-thickness
 * path of bone-morphometry 0.7.0-rc.3.
 *
 * Each entry point replaces one function of the reference crate (paths relative
 * to the crate root) and keeps its argument meaning, its output layout (z-major,
 * index (z*height + y)*width + x, src/volume.rs:44-46) and its error behaviour:
 * every function returns BM_OK or an error code, and bm_last_error() returns the
 * reference's message text (the anyhow::bail! strings) for the calling thread.
 * Nothing unwinds across this boundary (the reference's release profile aborts
 * on panic, Cargo.toml:23).
 *
 * Host-pointer entry points copy the input to the current HIP device, compute
 * there, and copy the result back; outputs are caller-allocated.  The *_device
 * variants take device pointers and an optional hipStream_t (NULL = the
 * library's stream of that device) and leave the data resident.
 *
 * Thread safety: calls may come from any thread; work on one device is
 * serialised by the library.
 */
#ifndef BM_H
#define BM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_OK 0
#define BM_ERR_INVALID 1     /* input rejected (the reference's bail! cases) */
#define BM_ERR_DEVICE 2      /* HIP runtime error or device unavailable */
#define BM_ERR_UNSUPPORTED 3 /* outside this implementation's size limits */

/* Phase, src/edt.rs:8-24: FOREGROUND selects value != 0, BACKGROUND value == 0. */
#define BM_PHASE_FOREGROUND 0
#define BM_PHASE_BACKGROUND 1

/* FeatureMode, src/metrics.rs:12-24 (declaration order). */
#define BM_MODE_BONEJ 0
#define BM_MODE_STANDARD 1
#define BM_MODE_CLASSIC 2
#define BM_MODE_REFINED 3
#define BM_MODE_EXPERIMENTAL 4

/* BinaryVolume, src/volume.rs:3-9: data[len] with len == width*height*depth,
 * every dimension >= 2 (src/volume.rs:12-41). */
typedef struct bm_volume {
  const uint8_t *data;
  uint64_t len;
  uint64_t width, height, depth;
} bm_volume;

/* DistributionStats, src/metrics.rs:26-38 (plus the number of values). */
typedef struct bm_distribution_stats {
  double mean;
  double standard_deviation;
  double maximum;
  double median;
  double percentile_10;
  double percentile_25;
  double percentile_75;
  double percentile_90;
  double interquartile_range;
  double coefficient_of_variation;
  uint64_t count;
} bm_distribution_stats;

/* The thickness/spacing part of Metrics, src/metrics.rs:48-62. */
typedef struct bm_thickness_spacing {
  bm_distribution_stats thickness; /* Tb.Th, src/metrics.rs:111-117 */
  bm_distribution_stats spacing;   /* Tb.Sp, src/metrics.rs:119-121 */
  double bone_volume;              /* src/metrics.rs:123-132 */
  double total_volume;
  double bone_volume_fraction;
} bm_thickness_spacing;

/* Message of the last failed call on this thread ("" if none). */
const char *bm_last_error(void);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int bm_version(void);

/* Number of HIP devices visible to the process (0 without a GPU). */
int bm_device_count(int *count);

/* Replaces the checks of BinaryVolume::new, src/volume.rs:12-41: returns
 * BM_OK, or BM_ERR_INVALID with the reference's message. */
int bm_binary_volume_check(const bm_volume *volume);

/* Replaces BinaryVolume::phase_inside(&self, roi, select_bone),
 * src/volume.rs:84-105: out[i] = (roi[i] != 0 && (bone[i] != 0) == select_bone),
 * computed on the device.  Fails like the reference on a shape mismatch. */
int bm_phase_inside(const bm_volume *bone, const bm_volume *roi, int select_bone, uint8_t *out,
                    uint64_t out_len);

/* Replaces euclidean_distance_transform_squared(&BinaryVolume, Phase)
 * -> Result<Vec<u32>>, src/edt.rs:57-156.  out[out_len] with out_len == len.
 * Fails like the reference when every voxel is selected (src/edt.rs:61-66). */
int bm_euclidean_distance_transform_squared(const bm_volume *volume, int phase, uint32_t *out,
                                            uint64_t out_len);

/* Replaces local_thickness(&BinaryVolume, Phase) -> Result<Vec<f32>>,
 * src/local_thickness.rs:16-44.  out[out_len] receives diameters in voxel
 * units, 0 outside the selected phase. */
int bm_local_thickness(const bm_volume *volume, int phase, float *out, uint64_t out_len);

/* Replaces the thickness/spacing part of analyze(bone, roi, voxel_spacing,
 * feature_mode), src/metrics.rs:64-132: the validation of :70-104, the two
 * ROI-clipped phases of :97-98 (src/volume.rs:84-105), local thickness of
 * each and nonzero_stats (src/metrics.rs:204-269) scaled by voxel_spacing[0],
 * and BV, TV, BV/TV.  Quantiles are filled for STANDARD and EXPERIMENTAL,
 * NaN otherwise, as in the reference. */
int bm_analyze_thickness_spacing(const bm_volume *bone, const bm_volume *roi,
                                 const double voxel_spacing[3], int feature_mode,
                                 bm_thickness_spacing *out);

/* Order of the f64 sums behind the mean and standard deviation of the
 * thickness statistics (no counterpart in the reference, whose order is
 * BM_STATS_SEQUENTIAL: src/metrics.rs:215-226 sums the map in f64, value after
 * value, in linear index order).
 *   BM_STATS_EXACT (default): exact fixed-point sums on the device, rounded
 *     once to f64.  Mean and SD differ from the reference's only by the
 *     rounding of its sequential sums.
 *   BM_STATS_SEQUENTIAL: the finished maps are copied to the host and summed
 *     in the reference's order, so mean, SD and CV equal the reference's bit
 *     for bit (slower: one map copy and a serial pass per phase).
 * Process-wide; applies to the analyze calls (host and device forms, and
 * bm_analyze) and to batches created afterwards (the server and the
 * no-allocation entry point always use BM_STATS_EXACT).  *previous (may be
 * NULL) receives the order that was in force. */
#define BM_STATS_EXACT 0
#define BM_STATS_SEQUENTIAL 1
int bm_set_stats_order(int order, int *previous);

/* Device-resident forms of the three calls above: d_* are device pointers on
 * the current HIP device, stream is a hipStream_t or NULL.  Each returns once
 * its result is complete (the statistics are copied to *out). */
int bm_euclidean_distance_transform_squared_device(const uint8_t *d_data, uint64_t width,
                                                   uint64_t height, uint64_t depth, int phase,
                                                   uint32_t *d_out, void *stream);
int bm_local_thickness_device(const uint8_t *d_data, uint64_t width, uint64_t height,
                              uint64_t depth, int phase, float *d_out, void *stream);
int bm_analyze_thickness_spacing_device(const uint8_t *d_bone, const uint8_t *d_roi,
                                        uint64_t width, uint64_t height, uint64_t depth,
                                        const double voxel_spacing[3], int feature_mode,
                                        bm_thickness_spacing *out, void *stream);

/* No counterpart in the reference: bm_analyze_thickness_spacing_device for
 * callers that cannot allocate memory (code that must not call hipMalloc or
 * hipHostMalloc: tight serving loops, harnesses that forbid allocation).
 * Same arguments and results, bit for bit; every device buffer of the call
 * lives in a device array of BM_NOALLOC_SCRATCH_BYTES that the library holds
 * on each GPU it runs on until the process ends (about 9 million voxels, e.g.
 * 208^3, fit; larger volumes fail with BM_ERR_UNSUPPORTED before any device
 * work).  The statistics use BM_STATS_EXACT.  Calls on one GPU share that
 * array: the library serialises them, and each returns when its results are
 * complete. */
#define BM_NOALLOC_SCRATCH_BYTES (256ull << 20)
int bm_analyze_thickness_spacing_device_noalloc(const uint8_t *d_bone, const uint8_t *d_roi,
                                                uint64_t width, uint64_t height, uint64_t depth,
                                                const double voxel_spacing[3], int feature_mode,
                                                bm_thickness_spacing *out, void *stream);

/* Intermediate products of local_thickness for parity tests and diagnostics
 * (no counterpart in the reference's public API; these are the values of
 * src/local_thickness.rs:17,20-26,28 before the clean-up).  Any output pointer
 * may be NULL.  ridge_count/fill_visits receive the ridge count and the work
 * counter of the sphere fill. */
int bm_local_thickness_stages(const bm_volume *volume, int phase, uint32_t *squared_edt,
                              uint8_t *ridge_flags, uint32_t *propagated, float *thickness,
                              uint64_t *ridge_count, uint64_t *fill_visits);

/* ---- The rest of analyze: cubical measures, topology, BoneJ connectivity,
 * interface surface and the ROI-boundary fraction (src/topology.rs,
 * src/metrics.rs:134-202). */

/* CubicalMeasures, src/topology.rs:13-17. */
typedef struct bm_cubical {
  int64_t raw_euler;
  double mean_breadth_lattice_units;
} bm_cubical;

/* Topology, src/topology.rs:5-11. */
typedef struct bm_topology {
  int64_t raw_euler;
  uint64_t beta0, beta1, beta2;
} bm_topology;

/* BonejConnectivity, src/topology.rs:19-27. */
typedef struct bm_bonej {
  double edge_correction, delta_chi, connectivity;
} bm_bonej;

/* GeometryMetrics, src/metrics.rs:40-46. */
typedef struct bm_geometry {
  double bone_surface, bone_surface_to_bone_volume, bone_surface_to_total_volume, mean_breadth;
} bm_geometry;

/* Metrics, src/metrics.rs:48-62 (field order kept). */
typedef struct bm_metrics {
  bm_distribution_stats thickness, spacing;
  double bone_volume, total_volume, bone_volume_fraction;
  bm_topology topology;
  bm_bonej bonej_connectivity;
  double bonej_connectivity_density, connectivity, connectivity_density,
      roi_boundary_bone_fraction;
  bm_geometry geometry;
} bm_metrics;

/* Replaces cubical_measures(&BinaryVolume) -> CubicalMeasures,
 * src/topology.rs:53-171 (the volume's nonzero voxels are the complex). */
int bm_cubical_measures(const bm_volume *volume, bm_cubical *out);

/* Replaces analyze_topology(&BinaryVolume, raw_euler) -> Result<Topology>,
 * src/topology.rs:29-51: beta0 (26-connected components), beta2 (enclosed
 * 6-connected background components), beta1 = beta0 + beta2 - raw_euler, with
 * the reference's error when that is negative. */
int bm_analyze_topology(const bm_volume *volume, int64_t raw_euler, bm_topology *out);

/* Replaces bonej_connectivity(&BinaryVolume, raw_euler) -> BonejConnectivity,
 * src/topology.rs:175-201 (stack edge correction). */
int bm_bonej_connectivity(const bm_volume *volume, int64_t raw_euler, bm_bonej *out);

/* Replaces bone_marrow_interface_faces(&bone, &roi) -> Result<i64>,
 * src/topology.rs:442-490. */
int bm_bone_marrow_interface_faces(const bm_volume *bone, const bm_volume *roi, int64_t *faces);

/* Replaces analyze(bone, roi, voxel_spacing, feature_mode) -> Result<Metrics>,
 * src/metrics.rs:64-164, whole: thickness/spacing as
 * bm_analyze_thickness_spacing, then the measures above on the analysis
 * bone, in the reference's order of checks and errors. */
int bm_analyze(const bm_volume *bone, const bm_volume *roi, const double voxel_spacing[3],
               int feature_mode, bm_metrics *out);
int bm_analyze_device(const uint8_t *d_bone, const uint8_t *d_roi, uint64_t width,
                      uint64_t height, uint64_t depth, const double voxel_spacing[3],
                      int feature_mode, bm_metrics *out, void *stream);

/* ---- The signed distance field and its Gaussian smoothing, the EDT consumers
 * of the curvature and Minkowski measures (src/sdf_curvature.rs:75-105,
 * src/minkowski.rs:54-71).  Outputs are f64, z-major like the volumes. */

/* Replaces require_isotropic_spacing(spacing) -> Result<f64>,
 * src/sdf_curvature.rs:147-160: *h = spacing[0] or BM_ERR_INVALID with the
 * reference's message. */
int bm_require_isotropic_spacing(const double spacing[3], double *h);

/* Replaces gaussian_kernel(sigma) -> Vec<f64>, src/sdf_curvature.rs:182-198:
 * writes min(len, cap) weights and sets *len = 2r+1 (r = floor(4 sigma + 0.5);
 * 1 for sigma <= 0 or NaN).  BM_ERR_UNSUPPORTED beyond r = 2^20. */
int bm_gaussian_kernel(double sigma, double *out, uint64_t cap, uint64_t *len);

/* Replaces signed_distance_field_um(&BinaryVolume, h) -> Result<Vec<f64>>,
 * src/sdf_curvature.rs:162-180: +sqrt(d2) h on bone voxels, -sqrt(d2') h
 * elsewhere (d2, d2' the squared EDTs of the two phases).  The errors are the
 * EDT's: the foreground transform's first. */
int bm_signed_distance_field_um(const bm_volume *bone, double h, double *out, uint64_t out_len);

/* Replaces gaussian_smooth_nearest_in_place(&mut Vec<f64>, width, height,
 * depth, sigma_vox), src/sdf_curvature.rs:200-269: x, y then z passes with
 * clamped indices, in place.  data holds len == width*height*depth values. */
int bm_gaussian_smooth_nearest_in_place(double *data, uint64_t len, uint64_t width,
                                        uint64_t height, uint64_t depth, double sigma_vox);

/* The two calls above in sequence, as src/sdf_curvature.rs:98-105,
 * src/minkowski.rs:64-71 and src/sdf_curvature.rs:719-726 make them; the
 * unsmoothed field never leaves the device. */
int bm_smoothed_signed_distance_field_um(const bm_volume *bone, double h, double sigma_vox,
                                         double *out, uint64_t out_len);

/* Device-resident forms (device pointers; stream may be NULL).  They return
 * when the output is complete. */
int bm_signed_distance_field_um_device(const uint8_t *d_bone, uint64_t width, uint64_t height,
                                       uint64_t depth, double h, double *d_out, void *stream);
int bm_gaussian_smooth_nearest_in_place_device(double *d_data, uint64_t width, uint64_t height,
                                               uint64_t depth, double sigma_vox, void *stream);
int bm_smoothed_signed_distance_field_um_device(const uint8_t *d_bone, uint64_t width,
                                                uint64_t height, uint64_t depth, double h,
                                                double sigma_vox, double *d_out, void *stream);

/* ---- The input step: multi-page 8-bit TIFF volumes and their calibration
 * (src/tiff_io.rs).  path is a NUL-terminated file name; messages name it as
 * given.  Pages are decoded on `threads` host threads (0: every core). */

/* Replaces read_binary_tiff(path, threshold, strict_binary) -> Result<BinaryVolume>,
 * src/tiff_io.rs:18-26,33-147: dims receives width, height, depth.  With out
 * NULL only the page directory is read (to size the buffer); otherwise every
 * page is decoded, its samples checked (0 or 255 when strict_binary) and
 * written as value > threshold, 0/1, into out (capacity bytes). */
int bm_read_binary_tiff(const char *path, uint8_t threshold, int strict_binary, int threads,
                        uint8_t *out, uint64_t capacity, uint64_t dims[3]);

/* Replaces read_roi_tiff(path) -> Result<BinaryVolume>, src/tiff_io.rs:28-31:
 * samples 0/1 or 0/255, written as value > 0. */
int bm_read_roi_tiff(const char *path, int threads, uint8_t *out, uint64_t capacity,
                     uint64_t dims[3]);

/* Device forms: decoded pages stream through pinned staging buffers into
 * d_out (capacity bytes of device memory) on stream while later pages decode;
 * they return when the volume is resident. */
int bm_read_binary_tiff_device(const char *path, uint8_t threshold, int strict_binary,
                               int threads, uint8_t *d_out, uint64_t capacity, uint64_t dims[3],
                               void *stream);
int bm_read_roi_tiff_device(const char *path, int threads, uint8_t *d_out, uint64_t capacity,
                            uint64_t dims[3], void *stream);

/* Replaces read_isotropic_spacing_um(path) -> Result<[f64; 3]>,
 * src/tiff_io.rs:149-317: x/y from XResolution, YResolution and
 * ResolutionUnit (or the ImageJ unit=), z from the ImageJ spacing=, with the
 * reference's isotropy checks and messages. */
int bm_read_isotropic_spacing_um(const char *path, double spacing[3]);

/* ---- Batched form (no counterpart in the reference: many analyze() calls,
 * src/metrics.rs:64-132, in one launch sequence).  The volumes are
 * device-resident, of any shapes; each is analyzed exactly as by
 * bm_analyze_thickness_spacing_device (same results, bit for bit).
 * bm_batch_create lays out one plan for all of them (bone and marrow of every
 * volume are items of the same launches) and, with use_graph, captures the
 * launch sequence into a hipGraph once; bm_batch_run replays it and waits.
 * Inputs may change between runs (same pointers and shapes). */
typedef struct bm_batch_item {
  const uint8_t *d_bone; /* device pointers, width*height*depth bytes each */
  const uint8_t *d_roi;
  uint64_t width, height, depth;
  double voxel_spacing[3];
  int feature_mode;
} bm_batch_item;

/* Per-volume outcome: status is what bm_analyze_thickness_spacing_device
 * would return for the volume, message its bm_last_error() text. */
typedef struct bm_batch_result {
  bm_thickness_spacing metrics;
  int status;
  char message[252];
} bm_batch_result;

typedef struct bm_batch bm_batch;
int bm_batch_create(const bm_batch_item *items, uint64_t count, int use_graph, bm_batch **out);
int bm_batch_run(bm_batch *batch, bm_batch_result *results /* count entries */);
int bm_batch_device_bytes(const bm_batch *batch, uint64_t *bytes);
int bm_batch_destroy(bm_batch *batch);

/* Kernel tunables per shape class: bm_autotune times the candidates on the
 * largest volume of each class among items and writes the choices to the
 * JSON cache at cache_path (merged with what it holds); bm_tuning_load makes
 * a cache current.  BM_TUNE_CACHE names the cache read at first use. */
int bm_autotune(const bm_batch_item *items, uint64_t count, const char *cache_path, int verbose);
int bm_tuning_load(const char *cache_path);

/* ---- Serving form (no counterpart in the reference).  A server owns `slots`
 * device regions of budget_bytes/slots each.  bm_server_submit queues a
 * host-resident volume (bone and roi stay valid and unchanged until the
 * request is polled done) and returns a ticket; a worker thread admits queued
 * requests, first come first served, into the next batch of a free slot while
 * the batch's device bytes fit the slot and it holds at most max_batch
 * volumes, uploads them and launches the batch; batches of different slots
 * overlap on the device.  bm_server_poll returns the result once done (with
 * wait != 0 it blocks until then) and releases the request. */
typedef struct bm_request_timing {
  double submit_s; /* seconds since the server started */
  double start_s;  /* its batch was launched */
  double done_s;   /* its batch completed */
  uint32_t batch_size;
  uint32_t pad;
} bm_request_timing;

typedef struct bm_server_stats {
  uint64_t batches;
  uint64_t requests_done;
  uint64_t max_batch;
  uint64_t sum_batch;
  uint64_t peak_device_bytes; /* largest sum of in-flight batches' bytes */
  uint32_t max_inflight;      /* most batches in flight at once */
  uint32_t slots;
  double busy_s;              /* time with at least one batch in flight */
} bm_server_stats;

typedef struct bm_server bm_server;
int bm_server_create(uint64_t budget_bytes, int max_batch, int slots, bm_server **out);
int bm_server_submit(bm_server *server, const uint8_t *bone, const uint8_t *roi, uint64_t width,
                     uint64_t height, uint64_t depth, const double voxel_spacing[3],
                     int feature_mode, uint64_t *ticket);
int bm_server_poll(bm_server *server, uint64_t ticket, int wait, bm_batch_result *result,
                   bm_request_timing *timing, int *done);
/* Withdraws a request that is still queued: *cancelled = 1, and its poll then
 * returns BM_ERR_INVALID with "request cancelled"; a request already launched
 * or done is left alone (*cancelled = 0).  Other requests are unaffected (the
 * reference's per-file isolation, src/main.rs:254-268). */
int bm_server_cancel(bm_server *server, uint64_t ticket, int *cancelled);
int bm_server_stats_get(bm_server *server, bm_server_stats *out);
int bm_server_destroy(bm_server *server);

/* Kernel timing for benchmarks and diagnostics (no counterpart in the
 * reference).  bm_profile_enable(1) brackets every kernel launch with HIP
 * events on its stream; bm_profile_read() waits for them and returns, per
 * kernel name, the launch count and total milliseconds since the last read. */
typedef struct bm_kernel_time {
  char name[48];
  uint64_t launches;
  double total_ms;
} bm_kernel_time;
int bm_profile_enable(int on);
int bm_profile_read(bm_kernel_time *out, int max, int *count);
/* Restricts the timed launches to the kernels named in names (comma-separated,
 * the names bm_profile_read reports); null or "" times every kernel.  Fewer
 * events keep the timed region closer to an unprofiled run (no counterpart in
 * the reference). */
int bm_profile_only(const char *names);

#ifdef __cplusplus
}
#endif

#endif /* BM_H */
