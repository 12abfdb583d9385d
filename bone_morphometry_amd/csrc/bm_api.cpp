// bm_api.cpp -- the reference crate's interface for this path, in C++, over
// the device engine, and the C ABI of include/bm.h.
//
// Names, argument meaning and error messages follow bone-morphometry 0.7.0-rc.3:
//   BinaryVolume::new           src/volume.rs:12-41
//   Phase                       src/edt.rs:8-32
//   euclidean_distance_transform_squared   src/edt.rs:57-156
//   local_thickness             src/local_thickness.rs:16-44
//   analyze (thickness/spacing part)       src/metrics.rs:64-132
//   nonzero_stats               src/metrics.rs:204-269
// The arithmetic of the statistics runs here on the host from exact sums
// computed on the device; this file must be built with -ffp-contract=off so
// count*sum_of_squares - sum*sum is evaluated unfused, as Rust does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/bm.h"
#include "bm_analyze.h"
#include "bm_engine.h"
#include "bm_serve.h"
#include "bm_smooth.h"
#include "bm_tiff.h"
#include "bm_topology.h"

namespace bm {

enum class Phase { Foreground, Background };

static const char* phase_name(Phase p) {
  return p == Phase::Foreground ? "foreground" : "background";
}

static std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
static std::string fmt(const char* f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// BinaryVolume (src/volume.rs:3-9) as a borrowed view; create() applies the
// checks of BinaryVolume::new (src/volume.rs:12-41).
struct BinaryVolume {
  const uint8_t* data = nullptr;
  uint64_t width = 0, height = 0, depth = 0;

  static BinaryVolume create(const uint8_t* data, uint64_t len, uint64_t w, uint64_t h,
                             uint64_t d) {
    unsigned __int128 expected = (unsigned __int128)w * h * d;
    if (expected > (unsigned __int128)SIZE_MAX) throw Error("volume dimensions overflow usize");
    if ((unsigned __int128)len != expected)
      throw Error(fmt("volume has %llu voxels, but dimensions %llux%llux%llu require %llu",
                      (unsigned long long)len, (unsigned long long)w, (unsigned long long)h,
                      (unsigned long long)d, (unsigned long long)expected));
    if (w < 2 || h < 2 || d < 2)
      throw Error(fmt("all dimensions must be at least 2 for 3-D morphometry; got %llux%llux%llu",
                      (unsigned long long)w, (unsigned long long)h, (unsigned long long)d));
    BinaryVolume v;
    v.data = data;
    v.width = w;
    v.height = h;
    v.depth = d;
    return v;
  }
  uint64_t len() const { return width * height * depth; }
  bool same_shape(const BinaryVolume& o) const {
    return width == o.width && height == o.height && depth == o.depth;
  }
};

// src/volume.rs:86-94 and src/metrics.rs:79-89
static std::string shape_mismatch(const BinaryVolume& a, const BinaryVolume& b) {
  return fmt("binary image dimensions %llux%llux%llu do not match ROI dimensions %llux%llux%llu",
             (unsigned long long)a.width, (unsigned long long)a.height,
             (unsigned long long)a.depth, (unsigned long long)b.width,
             (unsigned long long)b.height, (unsigned long long)b.depth);
}

// Limits of the device path (u16 row distances, u32 voxel indices).
Dims device_dims(uint64_t w, uint64_t h, uint64_t d) {
  if (w < 2 || h < 2 || d < 2)
    throw Error(fmt("all dimensions must be at least 2 for 3-D morphometry; got %llux%llux%llu",
                    (unsigned long long)w, (unsigned long long)h, (unsigned long long)d));
  if (w > 16384 || h > 16384 || d > 16384 || (unsigned __int128)w * h * d > 0xFFFFFFFFull)
    throw std::range_error(
        fmt("volume %llux%llux%llu exceeds the device path's limits (each axis <= 16384, "
            "fewer than 2^32 voxels)",
            (unsigned long long)w, (unsigned long long)h, (unsigned long long)d));
  Dims dims;
  dims.w = (uint32_t)w;
  dims.h = (uint32_t)h;
  dims.d = (uint32_t)d;
  return dims;
}

// Kernel tunables for a volume: the autotuner's choice for its shape class,
// with test overrides from the environment (BM_FILL_CAP: fill candidates per
// round; BM_FILL_TILES: tiles per fill block; BM_TILE_ORDER: tile order bits;
// BM_FILL_DOT: 0 for the packed-position candidate test at every radius;
// BM_FILL_CHUNK_MAP, BM_FILL_STAGE: 0 to turn off the fill's chunk descriptors or
// its staging area, so that the tests reach the paths large radii take;
// BM_EDT_TILES "y_rows,y_halo,z_rows,z_halo": the Y and Z passes' tiles).
static Tuning tuning_for(const Dims& dims) {
  Tuning t = tuned_for(dims);
  if (const char* env = std::getenv("BM_FILL_CAP")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v >= 1 && v <= 1024) t.fill_cap = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_FILL_TILES")) {  // tiles per fill block (1 or 2)
    const long v = std::strtol(env, nullptr, 10);
    if (v == 1 || v == 2) t.fill_tiles = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_FILL_DOT")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v == 0 || v == 1) t.fill_dot = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_FILL_CHUNK_MAP")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v == 0 || v == 1) t.fill_chunk_map = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_FILL_STAGE")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v == 0 || v == 1) t.fill_stage = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_TILE_ORDER")) {  // tile order bits (see Tuning)
    const long v = std::strtol(env, nullptr, 10);
    if (v >= 0 && v <= 7) t.tile_order = (uint32_t)v;
  }
  if (const char* env = std::getenv("BM_EDT_TILES")) {
    unsigned v[4];
    if (std::sscanf(env, "%u,%u,%u,%u", &v[0], &v[1], &v[2], &v[3]) == 4 && v[0] >= 2 &&
        v[0] <= 512 && v[2] >= 2 && v[2] <= 512) {
      t.y_rows = v[0];
      t.y_halo = v[1];
      t.z_rows = v[2];
      t.z_halo = v[3];
    }
  }
  return t;
}

static bool fill_diag_requested() { return std::getenv("BM_FILL_DIAG") != nullptr; }

// f64 in the shortest form that reads back exactly, like Rust's {:?}.
static std::string rust_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  for (int prec = 1; prec <= 17; ++prec) {
    snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s = buf;
  size_t e = s.find('e');
  if (e != std::string::npos) {
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);
    if (!ex.empty() && ex[0] == '+') ex = ex.substr(1);
    size_t nz = 0;
    bool neg = !ex.empty() && ex[0] == '-';
    std::string digits = neg ? ex.substr(1) : ex;
    while (nz + 1 < digits.size() && digits[nz] == '0') ++nz;
    return mant + "e" + (neg ? "-" : "") + digits.substr(nz);
  }
  if (s.find('.') == std::string::npos) s += ".0";
  return s;
}

static std::string spacing_list(const double sp[3]) {
  return "[" + rust_f64(sp[0]) + ", " + rust_f64(sp[1]) + ", " + rust_f64(sp[2]) + "]";
}

// A ridge radius the tile lists cannot hold (status error bit 0): outside the
// device path's limits.
static void fail_list_range() {
  throw std::range_error(
      "a squared distance exceeds the device path's limit for ridge radii (r^2 < 2^23)");
}

static void fail_all_selected(Phase p) {
  throw Error(std::string("distance transform is undefined for an all-") + phase_name(p) +
              " volume unless an exterior convention is specified");
}

// ---------------------------------------------------------------------------
// euclidean_distance_transform_squared (src/edt.rs:57-156), device resident:
// a plan of one item without the thickness stages.  d_out may be null (the
// result stays in the plan's buffer).
static const Plan& edt_device(Engine& e, const uint8_t* d_data, const Dims& dims, Phase p,
                              uint32_t* d_out, hipStream_t s) {
  ItemSpec sp;
  sp.src.data = d_data;
  sp.src.want_nonzero = p == Phase::Foreground;
  sp.dims = dims;
  sp.tune = tuning_for(dims);
  sp.d2_out = d_out;
  sp.thickness = false;
  Plan& pl = e.plain_plan();
  pl.build({sp}, s);
  pl.launch(s);
  pl.fetch(s);
  if (!pl.result(0).status.has_site) fail_all_selected(p);
  return pl;
}

void euclidean_distance_transform_squared(const BinaryVolume& v, Phase p, uint32_t* out) {
  const Dims dims = device_dims(v.width, v.height, v.depth);
  Engine& e = engine_for_device(current_device());
  std::lock_guard<std::mutex> g(e.mutex());
  hipStream_t s = e.stream();
  const uint64_t n = dims.n();
  uint8_t* d_in = static_cast<uint8_t*>(e.scratch_bytes(n, 0));
  BM_HIP(hipMemcpyAsync(d_in, v.data, n, hipMemcpyHostToDevice, s));
  const Plan& pl = edt_device(e, d_in, dims, p, nullptr, s);
  BM_HIP(hipMemcpyAsync(out, pl.desc(0).d2, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  BM_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// local_thickness (src/local_thickness.rs:16-44), device resident: a plan of
// one item.  With stages, the ridge flags and the propagated map are written
// too and the status is read back after them (ridge count).
struct StageOutputs {
  uint8_t* ridge_flags = nullptr;
  uint32_t* propagated = nullptr;
  PhaseStatus status{};
};

static const Plan& local_thickness_device(Engine& e, const uint8_t* d_data, const Dims& dims,
                                          Phase p, float* d_out, hipStream_t s,
                                          StageOutputs* stages = nullptr) {
  ItemSpec sp;
  sp.src.data = d_data;
  sp.src.want_nonzero = p == Phase::Foreground;
  sp.dims = dims;
  sp.tune = tuning_for(dims);
  sp.map_out = d_out;
  sp.want_map = true;
  sp.thickness = true;
  sp.diag = stages && fill_diag_requested();
  sp.scale = 1.0f;
  sp.lsb = stats_lsb_exp(1.0f);
  Plan& pl = e.plain_plan();
  pl.build({sp}, s);
  pl.launch(s);
  if (stages) e.stage_outputs(pl, stages->ridge_flags, stages->propagated, s);
  pl.fetch(s);
  if (!pl.result(0).status.has_site) fail_all_selected(p);
  if (pl.result(0).status.error) fail_list_range();
  if (stages) {
    BM_HIP(hipMemcpyAsync(&stages->status, pl.desc(0).status, sizeof(PhaseStatus),
                          hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
  }
  return pl;
}

void local_thickness(const BinaryVolume& v, Phase p, float* out) {
  const Dims dims = device_dims(v.width, v.height, v.depth);
  Engine& e = engine_for_device(current_device());
  std::lock_guard<std::mutex> g(e.mutex());
  hipStream_t s = e.stream();
  const uint64_t n = dims.n();
  char* stage = static_cast<char*>(e.scratch_bytes(n * 5 + 256, 0));
  uint8_t* d_in = reinterpret_cast<uint8_t*>(stage);
  float* d_out = reinterpret_cast<float*>(stage + ((n + 255) / 256) * 256);
  BM_HIP(hipMemcpyAsync(d_in, v.data, n, hipMemcpyHostToDevice, s));
  local_thickness_device(e, d_in, dims, p, d_out, s);
  BM_HIP(hipMemcpyAsync(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
  BM_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// nonzero_stats (src/metrics.rs:204-269) from the exact device sums.
static bm_distribution_stats distribution_stats(const StatsResult& r, const QuantileResult* q) {
  if (r.flags) throw DeviceError("statistics accumulator left its exact range");
  if (r.count == 0) throw Error("cannot compute statistics because the selected phase is empty");
  const double count_f64 = (double)r.count;
  const double mean = r.sum / count_f64;
  const double centered = (count_f64 * r.sum_of_squares - r.sum * r.sum) / count_f64;
  const double sd =
      (r.count > 1 && centered > 0.0) ? std::sqrt(centered / (count_f64 - 1.0)) : 0.0;
  bm_distribution_stats st;
  st.mean = mean;
  st.standard_deviation = sd;
  st.maximum = r.maximum;
  st.median = q ? q->median : NAN;
  st.percentile_10 = q ? q->p10 : NAN;
  st.percentile_25 = q ? q->p25 : NAN;
  st.percentile_75 = q ? q->p75 : NAN;
  st.percentile_90 = q ? q->p90 : NAN;
  st.interquartile_range = st.percentile_75 - st.percentile_25;
  st.coefficient_of_variation = sd / mean;
  st.count = r.count;
  return st;
}

// Thickness/spacing part of analyze (src/metrics.rs:64-132) on device-resident
// bone and ROI.  Both phases are derived on the fly from (bone, roi)
// (src/volume.rs:101); the map of a phase is only materialised when quantiles
// are requested.
void check_spacing(const double spacing[3]) {
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(spacing[i]) || spacing[i] <= 0.0)
      throw Error("voxel spacings must be finite and positive; got " + spacing_list(spacing));
}

void check_mode(int mode) {
  if (mode < BM_MODE_BONEJ || mode > BM_MODE_EXPERIMENTAL)
    throw Error(fmt("unknown feature mode %d", mode));
}

// The two items (bone, marrow) of one analyze call.
void analyze_items(const uint8_t* d_bone, const uint8_t* d_roi, const Dims& dims,
                   const double spacing[3], int mode, std::vector<ItemSpec>& out,
                   bool keep_maps) {
  const bool quantiles = mode == BM_MODE_STANDARD || mode == BM_MODE_EXPERIMENTAL;
  const float scale_f32 = (float)spacing[0];
  for (int ph = 0; ph < 2; ++ph) {
    ItemSpec sp;
    sp.src.data = d_bone;
    sp.src.roi = d_roi;
    sp.src.want_nonzero = ph == 0 ? 1u : 0u;
    sp.dims = dims;
    sp.tune = tuning_for(dims);
    sp.thickness = true;
    sp.quantiles = quantiles;
    sp.want_map = quantiles || keep_maps;
    sp.keep_d2 = false;
    sp.scale = scale_f32;
    sp.lsb = stats_lsb_exp(scale_f32);
    out.push_back(sp);
  }
}

// Results of analyze from the two items' results, with the error precedence
// of src/metrics.rs:90-104.
void analyze_finish(const ItemResult& bone, const ItemResult& marrow, uint64_t n,
                    const double spacing[3], int mode, bm_thickness_spacing* out,
                    const StatsResult* seq) {
  const bool quantiles = mode == BM_MODE_STANDARD || mode == BM_MODE_EXPERIMENTAL;
  const float scale_f32 = (float)spacing[0];
  const int lsb = stats_lsb_exp(scale_f32);
  const uint64_t bone_voxels = bone.status.selected, marrow_voxels = marrow.status.selected;
  const uint64_t roi_voxels = bone_voxels + marrow_voxels;
  if (roi_voxels == 0) throw Error("ROI mask contains no included voxels");
  if (mode == BM_MODE_BONEJ && roi_voxels == n)
    throw Error("BoneJ compatibility mode requires an ROI mask with both inside and outside voxels");
  if (bone_voxels == 0) throw Error("ROI contains no bone voxels");
  if (marrow_voxels == 0) throw Error("ROI contains no marrow voxels");
  if ((bone.status.error | marrow.status.error) != 0) fail_list_range();
  QuantileResult q[2];
  if (quantiles) {
    q[0] = quantiles_from(bone.sel, bone.stats.count, scale_f32);
    q[1] = quantiles_from(marrow.sel, marrow.stats.count, scale_f32);
  }
  StatsResult sums[2] = {finish_stats(bone.stats, lsb), finish_stats(marrow.stats, lsb)};
  if (seq) {
    for (int i = 0; i < 2; ++i) {
      if (sums[i].flags) throw DeviceError("statistics accumulator left its exact range");
      if (seq[i].count != sums[i].count)
        throw DeviceError("sequential statistics disagree with the device count");
      sums[i] = seq[i];
    }
  }
  out->thickness = distribution_stats(sums[0], quantiles ? &q[0] : nullptr);
  out->spacing = distribution_stats(sums[1], quantiles ? &q[1] : nullptr);
  const double voxel_volume = spacing[0] * spacing[1] * spacing[2];
  out->bone_volume = (double)bone_voxels * voxel_volume;
  out->total_volume = (double)roi_voxels * voxel_volume;
  out->bone_volume_fraction = out->bone_volume / out->total_volume;
}

// ---------------------------------------------------------------------------
// Order of the statistics sums (bm_set_stats_order).
static std::atomic<int> g_stats_order{BM_STATS_EXACT};

bool stats_sequential() { return g_stats_order.load() == BM_STATS_SEQUENTIAL; }

// The reference's loop (src/metrics.rs:215-226) over a device map: chunks come
// back through two pinned buffers, the next one in flight while the current
// one is summed, one value at a time in index order (this file is built with
// FMA contraction off, so sum_of_squares += value * value rounds twice).
StatsResult sequential_sums(const float* d_map, uint64_t n, float scale_f32, hipStream_t s) {
  constexpr size_t kChunk = (size_t)4 << 20;  // values per chunk (16 MiB)
  static std::mutex mu;
  static float* buf[2] = {nullptr, nullptr};
  std::lock_guard<std::mutex> g(mu);
  if (!buf[0]) {
    BM_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf[0]), kChunk * sizeof(float)));
    BM_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf[1]), kChunk * sizeof(float)));
  }
  uint64_t count = 0;
  double sum = 0.0, sum_of_squares = 0.0, maximum = 0.0;
  const uint64_t chunks = (n + kChunk - 1) / kChunk;
  auto issue = [&](uint64_t k) {
    const uint64_t off = k * kChunk, len = std::min<uint64_t>(kChunk, n - off);
    BM_HIP(hipMemcpyAsync(buf[k & 1], d_map + off, len * sizeof(float), hipMemcpyDeviceToHost,
                          s));
  };
  if (chunks) issue(0);
  BM_HIP(hipStreamSynchronize(s));
  for (uint64_t k = 0; k < chunks; ++k) {
    if (k + 1 < chunks) issue(k + 1);
    const float* b = buf[k & 1];
    const uint64_t len = std::min<uint64_t>(kChunk, n - k * kChunk);
    for (uint64_t i = 0; i < len; ++i) {
      const float raw = b[i];
      if (!std::isfinite(raw) || raw <= 0.0f) continue;
      const double value = (double)(raw * scale_f32);
      count += 1;
      sum += value;
      sum_of_squares += value * value;
      maximum = std::max(maximum, value);
    }
    BM_HIP(hipStreamSynchronize(s));
  }
  StatsResult r;
  r.count = count;
  r.sum = sum;
  r.sum_of_squares = sum_of_squares;
  r.maximum = maximum;
  return r;
}

bool replay_sums(const Plan& plan, size_t first, float scale_f32, hipStream_t s,
                 StatsResult out[2]) {
  for (int i = 0; i < 2; ++i) {
    const ItemResult& r = plan.result(first + i);
    if (r.status.selected == 0 || !r.status.has_site) return false;
  }
  for (int i = 0; i < 2; ++i) {
    const ItemDesc& d = plan.desc(first + i);
    out[i] = sequential_sums(d.map, d.n, scale_f32, s);
  }
  return true;
}

// Thickness/spacing part of analyze (src/metrics.rs:64-132) on device-resident
// bone and ROI: the bone and marrow phases (src/volume.rs:101) are two items
// of one plan, so both run in the same launches.
static void analyze_device(Engine& e, const uint8_t* d_bone, const uint8_t* d_roi,
                           const Dims& dims, const double spacing[3], int mode,
                           bm_thickness_spacing* out, hipStream_t s) {
  check_spacing(spacing);
  check_mode(mode);
  std::vector<ItemSpec> specs;
  const bool seq = stats_sequential();
  analyze_items(d_bone, d_roi, dims, spacing, mode, specs, seq);
  Plan& pl = e.plain_plan();
  pl.build(specs, s);
  pl.launch(s);
  pl.fetch(s);
  StatsResult sums[2];
  const bool replayed = seq && replay_sums(pl, 0, (float)spacing[0], s, sums);
  analyze_finish(pl.result(0), pl.result(1), dims.n(), spacing, mode, out,
                 replayed ? sums : nullptr);
}

void analyze_thickness_spacing(const BinaryVolume& bone, const BinaryVolume& roi,
                               const double spacing[3], int mode, bm_thickness_spacing* out) {
  check_spacing(spacing);
  if (!bone.same_shape(roi)) throw Error(shape_mismatch(bone, roi));
  const Dims dims = device_dims(bone.width, bone.height, bone.depth);
  Engine& e = engine_for_device(current_device());
  std::lock_guard<std::mutex> g(e.mutex());
  hipStream_t s = e.stream();
  const uint64_t n = dims.n();
  uint8_t* d_in = static_cast<uint8_t*>(e.scratch_bytes(2 * n + 256, 0));
  uint8_t* d_roi = d_in + ((n + 255) / 256) * 256;
  BM_HIP(hipMemcpyAsync(d_in, bone.data, n, hipMemcpyHostToDevice, s));
  BM_HIP(hipMemcpyAsync(d_roi, roi.data, n, hipMemcpyHostToDevice, s));
  analyze_device(e, d_in, d_roi, dims, spacing, mode, out, s);
}

// ---------------------------------------------------------------------------
// The rest of analyze (src/metrics.rs:134-164) from the device's integer counts.

// cubical_measures' two outputs, src/topology.rs:167-170.
static bm_cubical cubical_from(const TopologyCounts& t) {
  bm_cubical c;
  c.raw_euler = t.vertices - t.edges + t.faces - t.cubes;
  c.mean_breadth_lattice_units = 0.5 * (double)(t.edges - 2 * t.faces + 3 * t.cubes);
  return c;
}

// analyze_topology, src/topology.rs:29-51.
static bm_topology topology_from(const TopologyCounts& t, int64_t raw_euler) {
  const int64_t beta1 = (int64_t)t.beta0 + (int64_t)t.beta2 - raw_euler;
  if (beta1 < 0)
    throw Error(fmt("computed a negative beta1 (%lld); check the 26/6 topology convention",
                    (long long)beta1));
  bm_topology r;
  r.raw_euler = raw_euler;
  r.beta0 = t.beta0;
  r.beta1 = (uint64_t)beta1;
  r.beta2 = t.beta2;
  return r;
}

// bonej_connectivity, src/topology.rs:181-201.
static bm_bonej bonej_from(const TopologyCounts& t, int64_t raw_euler) {
  const int64_t f = t.stack_vertices;
  const int64_t e = t.stack_edges + 3 * f;
  const int64_t c = t.stack_faces + 2 * e - 3 * f;
  const int64_t d = t.edge_vertices + f;
  const int64_t a = t.face_vertices;
  const int64_t b = t.face_edges;
  const double chi_zero = (double)f;
  const double chi_one = (double)(d - e);
  const double chi_two = (double)(a - b + c);
  bm_bonej r;
  r.edge_correction = chi_two / 2.0 + chi_one / 4.0 + chi_zero / 8.0;
  r.delta_chi = (double)raw_euler - r.edge_correction;
  r.connectivity = 1.0 - r.delta_chi;
  return r;
}

// analyze, src/metrics.rs:64-164, on device-resident bone and ROI: the
// thickness plan and the topology passes run back to back on one stream;
// errors come in the reference's order (the thickness part's first).
static void analyze_full_device(Engine& e, const uint8_t* d_bone, const uint8_t* d_roi,
                                const Dims& dims, const double spacing[3], int mode,
                                bm_metrics* out, hipStream_t s) {
  check_spacing(spacing);
  check_mode(mode);
  std::vector<ItemSpec> specs;
  const bool seq = stats_sequential();
  analyze_items(d_bone, d_roi, dims, spacing, mode, specs, seq);
  Plan& pl = e.plain_plan();
  pl.build(specs, s);
  pl.launch(s);
  TopologyCounts t;
  topology_device(e, d_bone, d_roi, dims, kTopoCubical | kTopoBonej | kTopoBetti, &t, s);
  pl.fetch(s);
  StatsResult sums[2];
  const bool replayed = seq && replay_sums(pl, 0, (float)spacing[0], s, sums);
  bm_thickness_spacing ts;
  analyze_finish(pl.result(0), pl.result(1), dims.n(), spacing, mode, &ts,
                 replayed ? sums : nullptr);
  out->thickness = ts.thickness;
  out->spacing = ts.spacing;
  out->bone_volume = ts.bone_volume;
  out->total_volume = ts.total_volume;
  out->bone_volume_fraction = ts.bone_volume_fraction;
  const bm_cubical cubical = cubical_from(t);
  out->topology = topology_from(t, cubical.raw_euler);
  out->bonej_connectivity = bonej_from(t, cubical.raw_euler);
  out->bonej_connectivity_density = out->bonej_connectivity.connectivity / out->total_volume;
  out->connectivity = (double)out->topology.beta1;
  out->connectivity_density = out->connectivity / out->total_volume;
  const double bone_surface = (double)t.interface_faces * spacing[0] * spacing[1];
  out->geometry.bone_surface = bone_surface;
  out->geometry.bone_surface_to_bone_volume = bone_surface / out->bone_volume;
  out->geometry.bone_surface_to_total_volume = bone_surface / out->total_volume;
  out->geometry.mean_breadth = cubical.mean_breadth_lattice_units * spacing[0];
  out->roi_boundary_bone_fraction = (double)t.boundary_bone_voxels / (double)t.bone_voxels;
}

// One volume's counts (the nonzero voxels, no ROI), uploaded from the host.
static TopologyCounts volume_counts(const BinaryVolume& v, unsigned parts) {
  const Dims dims = device_dims(v.width, v.height, v.depth);
  Engine& e = engine_for_device(current_device());
  std::lock_guard<std::mutex> g(e.mutex());
  hipStream_t s = e.stream();
  uint8_t* d = static_cast<uint8_t*>(e.scratch_bytes(dims.n(), 0));
  BM_HIP(hipMemcpyAsync(d, v.data, dims.n(), hipMemcpyHostToDevice, s));
  TopologyCounts t;
  topology_device(e, d, nullptr, dims, parts, &t, s);
  return t;
}

// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

template <class F>
static int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return BM_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return BM_ERR_INVALID;
  } catch (const std::range_error& e) {
    g_last_error = e.what();
    return BM_ERR_UNSUPPORTED;
  } catch (const DeviceError& e) {
    g_last_error = e.what();
    return BM_ERR_DEVICE;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BM_ERR_DEVICE;
  } catch (...) {
    g_last_error = "unknown error";
    return BM_ERR_DEVICE;
  }
}

int capture_errors(void (*f)(void*), void* ctx, std::string* message) {
  const int rc = guarded([&] { f(ctx); });
  if (message) *message = g_last_error;
  return rc;
}

static BinaryVolume from_c(const bm_volume* v) {
  if (!v || (!v->data && v->len)) throw Error("null volume");
  return BinaryVolume::create(v->data, v->len, v->width, v->height, v->depth);
}

static Phase phase_from_c(int phase) {
  if (phase == BM_PHASE_FOREGROUND) return Phase::Foreground;
  if (phase == BM_PHASE_BACKGROUND) return Phase::Background;
  throw Error(fmt("unknown phase %d", phase));
}

static void check_out_len(uint64_t out_len, uint64_t n) {
  if (out_len != n)
    throw Error(fmt("output buffer holds %llu values, the volume has %llu voxels",
                    (unsigned long long)out_len, (unsigned long long)n));
}

static hipStream_t stream_or_default(Engine& e, void* stream) {
  return stream ? static_cast<hipStream_t>(stream) : e.stream();
}

}  // namespace bm

using namespace bm;

extern "C" {

const char* bm_last_error(void) { return g_last_error.c_str(); }

int bm_version(void) { return 100; }

int bm_profile_enable(int on) {
  return guarded([&] { profile_enable(on != 0); });
}

int bm_profile_only(const char* names) {
  return guarded([&] { profile_only(names); });
}

int bm_profile_read(bm_kernel_time* out, int max, int* count) {
  return guarded([&] {
    static_assert(sizeof(bm_kernel_time) == sizeof(KernelTime), "layout");
    int n = profile_read(reinterpret_cast<KernelTime*>(out), max);
    if (count) *count = n;
  });
}

int bm_device_count(int* count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  if (count) *count = c;
  return BM_OK;
}

int bm_binary_volume_check(const bm_volume* volume) {
  return guarded([&] { (void)from_c(volume); });
}

int bm_phase_inside(const bm_volume* bone, const bm_volume* roi, int select_bone, uint8_t* out,
                    uint64_t out_len) {
  return guarded([&] {
    BinaryVolume b = from_c(bone);
    BinaryVolume r = from_c(roi);
    if (!b.same_shape(r)) throw Error(shape_mismatch(b, r));
    check_out_len(out_len, b.len());
    const uint64_t n = b.len();
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    uint8_t* d = static_cast<uint8_t*>(e.scratch_bytes(3 * n + 512, 0));
    uint8_t* d_roi = d + ((n + 255) / 256) * 256;
    uint8_t* d_out = d_roi + ((n + 255) / 256) * 256;
    BM_HIP(hipMemcpyAsync(d, b.data, n, hipMemcpyHostToDevice, s));
    BM_HIP(hipMemcpyAsync(d_roi, r.data, n, hipMemcpyHostToDevice, s));
    e.run_phase_inside(d, d_roi, n, select_bone != 0, d_out, s);
    BM_HIP(hipMemcpyAsync(out, d_out, n, hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
  });
}

int bm_euclidean_distance_transform_squared(const bm_volume* volume, int phase, uint32_t* out,
                                            uint64_t out_len) {
  return guarded([&] {
    BinaryVolume v = from_c(volume);
    Phase p = phase_from_c(phase);
    check_out_len(out_len, v.len());
    euclidean_distance_transform_squared(v, p, out);
  });
}

int bm_local_thickness(const bm_volume* volume, int phase, float* out, uint64_t out_len) {
  return guarded([&] {
    BinaryVolume v = from_c(volume);
    Phase p = phase_from_c(phase);
    check_out_len(out_len, v.len());
    local_thickness(v, p, out);
  });
}

int bm_analyze_thickness_spacing(const bm_volume* bone, const bm_volume* roi,
                                 const double voxel_spacing[3], int feature_mode,
                                 bm_thickness_spacing* out) {
  return guarded([&] {
    BinaryVolume b = from_c(bone);
    BinaryVolume r = from_c(roi);
    analyze_thickness_spacing(b, r, voxel_spacing, feature_mode, out);
  });
}

int bm_euclidean_distance_transform_squared_device(const uint8_t* d_data, uint64_t width,
                                                   uint64_t height, uint64_t depth, int phase,
                                                   uint32_t* d_out, void* stream) {
  return guarded([&] {
    const Dims dims = device_dims(width, height, depth);
    Phase p = phase_from_c(phase);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    edt_device(e, d_data, dims, p, d_out, stream_or_default(e, stream));
  });
}

int bm_local_thickness_device(const uint8_t* d_data, uint64_t width, uint64_t height,
                              uint64_t depth, int phase, float* d_out, void* stream) {
  return guarded([&] {
    const Dims dims = device_dims(width, height, depth);
    Phase p = phase_from_c(phase);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    local_thickness_device(e, d_data, dims, p, d_out, stream_or_default(e, stream));
  });
}

int bm_set_stats_order(int order, int* previous) {
  return guarded([&] {
    if (order != BM_STATS_EXACT && order != BM_STATS_SEQUENTIAL)
      throw Error(fmt("unknown statistics order %d", order));
    const int prev = g_stats_order.exchange(order);
    if (previous) *previous = prev;
  });
}

int bm_analyze_thickness_spacing_device(const uint8_t* d_bone, const uint8_t* d_roi,
                                        uint64_t width, uint64_t height, uint64_t depth,
                                        const double voxel_spacing[3], int feature_mode,
                                        bm_thickness_spacing* out, void* stream) {
  return guarded([&] {
    const Dims dims = device_dims(width, height, depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    analyze_device(e, d_bone, d_roi, dims, voxel_spacing, feature_mode, out,
                   stream_or_default(e, stream));
  });
}

int bm_local_thickness_stages(const bm_volume* volume, int phase, uint32_t* squared_edt,
                              uint8_t* ridge_flags, uint32_t* propagated, float* thickness,
                              uint64_t* ridge_count, uint64_t* fill_visits) {
  return guarded([&] {
    BinaryVolume v = from_c(volume);
    Phase p = phase_from_c(phase);
    const Dims dims = device_dims(v.width, v.height, v.depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    const uint64_t n = dims.n();
    const uint64_t a = ((n + 255) / 256) * 256;
    char* stage = static_cast<char*>(e.scratch_bytes(a * 10, 0));
    uint8_t* d_in = reinterpret_cast<uint8_t*>(stage);
    uint8_t* d_flags = reinterpret_cast<uint8_t*>(stage + a);
    uint32_t* d_prop = reinterpret_cast<uint32_t*>(stage + 2 * a);
    float* d_map = reinterpret_cast<float*>(stage + 6 * a);
    BM_HIP(hipMemcpyAsync(d_in, v.data, n, hipMemcpyHostToDevice, s));
    StageOutputs so;
    so.ridge_flags = d_flags;
    so.propagated = d_prop;
    const Plan& pl = local_thickness_device(e, d_in, dims, p, d_map, s, &so);
    if (squared_edt)
      BM_HIP(hipMemcpyAsync(squared_edt, pl.desc(0).d2, n * 4, hipMemcpyDeviceToHost, s));
    if (ridge_flags) BM_HIP(hipMemcpyAsync(ridge_flags, d_flags, n, hipMemcpyDeviceToHost, s));
    if (propagated) BM_HIP(hipMemcpyAsync(propagated, d_prop, n * 4, hipMemcpyDeviceToHost, s));
    if (thickness) BM_HIP(hipMemcpyAsync(thickness, d_map, n * 4, hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
    if (ridge_count) *ridge_count = so.status.ridge_count;
    if (fill_visits) *fill_visits = so.status.fill_visits;
    if (std::getenv("BM_FILL_DIAG")) {
      const auto& d = so.status.fill_diag16;
      fprintf(stderr,
              "[fill] waves %llu tests %llu scanned %llu rounds %llu | clocks(10ns) collect %llu "
              "place %llu scan %llu total %llu | masks %llu nbrs %llu chunkscan %llu loads %llu "
              "rest %llu | kept %llu chunks %llu unstaged %llu overflow %u\n",
              d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12], d[13],
              d[14], d[15], so.status.overflow_tiles);
    }
  });
}

// ---- the input step: TIFF volumes ---------------------------------------------
static void tiff_dims(const TiffVolumeInfo& info, uint64_t dims[3]) {
  dims[0] = info.width;
  dims[1] = info.height;
  dims[2] = info.depth;
}

int bm_read_binary_tiff(const char* path, uint8_t threshold, int strict_binary, int threads,
                        uint8_t* out, uint64_t capacity, uint64_t dims[3]) {
  return guarded([&] {
    if (!path || !dims) throw Error("null argument");
    tiff_dims(read_binary_tiff(path, threshold,
                               strict_binary ? TiffSamples::ZeroOr255 : TiffSamples::Any, out,
                               capacity, threads),
              dims);
  });
}

int bm_read_roi_tiff(const char* path, int threads, uint8_t* out, uint64_t capacity,
                     uint64_t dims[3]) {
  return guarded([&] {
    if (!path || !dims) throw Error("null argument");
    tiff_dims(read_binary_tiff(path, 0, TiffSamples::RoiMask, out, capacity, threads), dims);
  });
}

int bm_read_binary_tiff_device(const char* path, uint8_t threshold, int strict_binary,
                               int threads, uint8_t* d_out, uint64_t capacity, uint64_t dims[3],
                               void* stream) {
  return guarded([&] {
    if (!path || !dims || !d_out) throw Error("null argument");
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    tiff_dims(read_binary_tiff_device(e, path, threshold,
                                      strict_binary ? TiffSamples::ZeroOr255 : TiffSamples::Any,
                                      d_out, capacity, threads, stream_or_default(e, stream)),
              dims);
  });
}

int bm_read_roi_tiff_device(const char* path, int threads, uint8_t* d_out, uint64_t capacity,
                            uint64_t dims[3], void* stream) {
  return guarded([&] {
    if (!path || !dims || !d_out) throw Error("null argument");
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    tiff_dims(read_binary_tiff_device(e, path, 0, TiffSamples::RoiMask, d_out, capacity, threads,
                                      stream_or_default(e, stream)),
              dims);
  });
}

int bm_read_isotropic_spacing_um(const char* path, double spacing[3]) {
  return guarded([&] {
    if (!path || !spacing) throw Error("null argument");
    read_isotropic_spacing_um(path, spacing);
  });
}

// ---- the rest of analyze -----------------------------------------------------
int bm_cubical_measures(const bm_volume* volume, bm_cubical* out) {
  return guarded([&] {
    if (!out) throw Error("null output");
    *out = cubical_from(volume_counts(from_c(volume), kTopoCubical));
  });
}

int bm_analyze_topology(const bm_volume* volume, int64_t raw_euler, bm_topology* out) {
  return guarded([&] {
    if (!out) throw Error("null output");
    *out = topology_from(volume_counts(from_c(volume), kTopoBetti), raw_euler);
  });
}

int bm_bonej_connectivity(const bm_volume* volume, int64_t raw_euler, bm_bonej* out) {
  return guarded([&] {
    if (!out) throw Error("null output");
    *out = bonej_from(volume_counts(from_c(volume), kTopoBonej), raw_euler);
  });
}

int bm_bone_marrow_interface_faces(const bm_volume* bone, const bm_volume* roi, int64_t* faces) {
  return guarded([&] {
    if (!faces) throw Error("null output");
    BinaryVolume b = from_c(bone);
    BinaryVolume r = from_c(roi);
    if (!b.same_shape(r))  // src/topology.rs:447-456
      throw Error(fmt("bone dimensions %llux%llux%llu do not match ROI dimensions %llux%llux%llu",
                      (unsigned long long)b.width, (unsigned long long)b.height,
                      (unsigned long long)b.depth, (unsigned long long)r.width,
                      (unsigned long long)r.height, (unsigned long long)r.depth));
    const Dims dims = device_dims(b.width, b.height, b.depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    const uint64_t n = dims.n();
    uint8_t* d = static_cast<uint8_t*>(e.scratch_bytes(2 * n + 256, 0));
    uint8_t* d_roi = d + ((n + 255) / 256) * 256;
    BM_HIP(hipMemcpyAsync(d, b.data, n, hipMemcpyHostToDevice, s));
    BM_HIP(hipMemcpyAsync(d_roi, r.data, n, hipMemcpyHostToDevice, s));
    // Only pairs with both voxels inside the ROI count, and there the bone
    // clipped to the ROI (what the lattice pass sees) is the bone as given.
    TopologyCounts t;
    topology_device(e, d, d_roi, dims, kTopoCubical, &t, s);
    *faces = t.interface_faces;
  });
}

int bm_analyze(const bm_volume* bone, const bm_volume* roi, const double voxel_spacing[3],
               int feature_mode, bm_metrics* out) {
  return guarded([&] {
    if (!out || !voxel_spacing) throw Error("null argument");
    BinaryVolume b = from_c(bone);
    BinaryVolume r = from_c(roi);
    check_spacing(voxel_spacing);
    if (!b.same_shape(r)) throw Error(shape_mismatch(b, r));
    const Dims dims = device_dims(b.width, b.height, b.depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    const uint64_t n = dims.n();
    uint8_t* d = static_cast<uint8_t*>(e.scratch_bytes(2 * n + 256, 0));
    uint8_t* d_roi = d + ((n + 255) / 256) * 256;
    BM_HIP(hipMemcpyAsync(d, b.data, n, hipMemcpyHostToDevice, s));
    BM_HIP(hipMemcpyAsync(d_roi, r.data, n, hipMemcpyHostToDevice, s));
    analyze_full_device(e, d, d_roi, dims, voxel_spacing, feature_mode, out, s);
  });
}

int bm_analyze_device(const uint8_t* d_bone, const uint8_t* d_roi, uint64_t width,
                      uint64_t height, uint64_t depth, const double voxel_spacing[3],
                      int feature_mode, bm_metrics* out, void* stream) {
  return guarded([&] {
    if (!out || !voxel_spacing) throw Error("null argument");
    const Dims dims = device_dims(width, height, depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    analyze_full_device(e, d_bone, d_roi, dims, voxel_spacing, feature_mode, out,
                        stream_or_default(e, stream));
  });
}

// ---- signed distance field and Gaussian smoothing ---------------------------
int bm_require_isotropic_spacing(const double spacing[3], double* h) {
  return guarded([&] {
    if (!spacing || !h) throw Error("null argument");
    *h = require_isotropic_spacing(spacing);
  });
}

int bm_gaussian_kernel(double sigma, double* out, uint64_t cap, uint64_t* len) {
  return guarded([&] {
    if (!len || (!out && cap)) throw Error("null argument");
    const std::vector<double> k = gaussian_kernel(sigma);
    *len = k.size();
    if (cap) std::memcpy(out, k.data(), std::min<uint64_t>(cap, k.size()) * sizeof(double));
  });
}

int bm_signed_distance_field_um(const bm_volume* bone, double h, double* out, uint64_t out_len) {
  return guarded([&] {
    BinaryVolume v = from_c(bone);
    const Dims dims = device_dims(v.width, v.height, v.depth);
    check_out_len(out_len, v.len());
    const uint64_t n = dims.n();
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    uint8_t* d_in = static_cast<uint8_t*>(e.scratch_bytes(n, 0));
    double* d_out = static_cast<double*>(e.scratch_bytes(n * sizeof(double), 2));
    BM_HIP(hipMemcpyAsync(d_in, v.data, n, hipMemcpyHostToDevice, s));
    signed_distance_field_device(e, d_in, dims, h, d_out, s);
    BM_HIP(hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
  });
}

int bm_gaussian_smooth_nearest_in_place(double* data, uint64_t len, uint64_t width,
                                        uint64_t height, uint64_t depth, double sigma_vox) {
  return guarded([&] {
    const Dims dims = smooth_dims(width, height, depth);
    if (len != dims.n())
      throw Error(fmt("data holds %llu values, the volume has %llu voxels",
                      (unsigned long long)len, (unsigned long long)dims.n()));
    if (!data && len) throw Error("null data");
    if (gaussian_kernel(sigma_vox).size() <= 1 || len == 0) return;
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    double* d = static_cast<double*>(e.scratch_bytes(len * sizeof(double), 0));
    BM_HIP(hipMemcpyAsync(d, data, len * sizeof(double), hipMemcpyHostToDevice, s));
    gaussian_smooth_device(e, d, dims, sigma_vox, s);
    BM_HIP(hipMemcpyAsync(data, d, len * sizeof(double), hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
  });
}

int bm_smoothed_signed_distance_field_um(const bm_volume* bone, double h, double sigma_vox,
                                         double* out, uint64_t out_len) {
  return guarded([&] {
    BinaryVolume v = from_c(bone);
    const Dims dims = device_dims(v.width, v.height, v.depth);
    check_out_len(out_len, v.len());
    const uint64_t n = dims.n();
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    hipStream_t s = e.stream();
    uint8_t* d_in = static_cast<uint8_t*>(e.scratch_bytes(n, 0));
    BM_HIP(hipMemcpyAsync(d_in, v.data, n, hipMemcpyHostToDevice, s));
    const double* d_out =
        smoothed_signed_distance_field_device(e, d_in, dims, h, sigma_vox, nullptr, s);
    BM_HIP(hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
  });
}

int bm_signed_distance_field_um_device(const uint8_t* d_bone, uint64_t width, uint64_t height,
                                       uint64_t depth, double h, double* d_out, void* stream) {
  return guarded([&] {
    const Dims dims = device_dims(width, height, depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    signed_distance_field_device(e, d_bone, dims, h, d_out, stream_or_default(e, stream));
  });
}

int bm_gaussian_smooth_nearest_in_place_device(double* d_data, uint64_t width, uint64_t height,
                                               uint64_t depth, double sigma_vox, void* stream) {
  return guarded([&] {
    const Dims dims = smooth_dims(width, height, depth);
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    gaussian_smooth_device(e, d_data, dims, sigma_vox, stream_or_default(e, stream));
  });
}

int bm_smoothed_signed_distance_field_um_device(const uint8_t* d_bone, uint64_t width,
                                                uint64_t height, uint64_t depth, double h,
                                                double sigma_vox, double* d_out, void* stream) {
  return guarded([&] {
    const Dims dims = device_dims(width, height, depth);
    if (!d_out) throw Error("null output");
    Engine& e = engine_for_device(current_device());
    std::lock_guard<std::mutex> g(e.mutex());
    smoothed_signed_distance_field_device(e, d_bone, dims, h, sigma_vox, d_out,
                                          stream_or_default(e, stream));
  });
}

// ---- batched form ----------------------------------------------------------
struct bm_batch {
  bm::Batch* impl;
};

static void to_c(const VolumeResult& v, bm_batch_result* out) {
  std::memset(out, 0, sizeof *out);
  out->metrics = v.metrics;
  out->status = v.status;
  std::strncpy(out->message, v.message.c_str(), sizeof out->message - 1);
}

int bm_batch_create(const bm_batch_item* items, uint64_t count, int use_graph, bm_batch** out) {
  return guarded([&] {
    if (!out) throw Error("null output");
    *out = nullptr;
    if (count && !items) throw Error("null items");
    if (count == 0) throw Error("a batch needs at least one volume");
    std::unique_ptr<bm_batch> b(new bm_batch{nullptr});
    b->impl = new Batch(items, count, use_graph != 0);
    *out = b.release();
  });
}

int bm_batch_run(bm_batch* batch, bm_batch_result* results) {
  return guarded([&] {
    if (!batch || !batch->impl || !results) throw Error("null batch or results");
    std::vector<VolumeResult> res;
    batch->impl->run(res);
    for (size_t i = 0; i < res.size(); ++i) to_c(res[i], &results[i]);
  });
}

int bm_batch_device_bytes(const bm_batch* batch, uint64_t* bytes) {
  return guarded([&] {
    if (!batch || !batch->impl || !bytes) throw Error("null batch or output");
    *bytes = batch->impl->device_bytes();
  });
}

int bm_batch_destroy(bm_batch* batch) {
  return guarded([&] {
    if (!batch) return;
    delete batch->impl;
    delete batch;
  });
}

int bm_autotune(const bm_batch_item* items, uint64_t count, const char* cache_path, int verbose) {
  return guarded([&] {
    if (!cache_path || (count && !items)) throw Error("null argument");
    std::map<std::string, TuneSample> largest;
    for (uint64_t i = 0; i < count; ++i) {
      const bm_batch_item& it = items[i];
      const Dims dims = device_dims(it.width, it.height, it.depth);
      const std::string cls = shape_class(dims);
      auto f = largest.find(cls);
      if (f == largest.end() || f->second.dims.n() < dims.n())
        largest[cls] = TuneSample{it.d_bone, it.d_roi, dims};
    }
    std::vector<TuneSample> samples;
    for (auto& kv : largest) samples.push_back(kv.second);
    autotune(samples, cache_path, verbose != 0);
  });
}

int bm_tuning_load(const char* cache_path) {
  return guarded([&] {
    if (!cache_path) throw Error("null path");
    tuning_load(cache_path);
  });
}

// ---- serving form -------------------------------------------------------------
struct bm_server {
  bm::Server* impl;
};

int bm_server_create(uint64_t budget_bytes, int max_batch, int slots, bm_server** out) {
  return guarded([&] {
    if (!out) throw Error("null output");
    *out = nullptr;
    if (budget_bytes == 0) throw Error("the HBM budget must be positive");
    std::unique_ptr<bm_server> s(new bm_server{nullptr});
    s->impl = new Server(budget_bytes, max_batch, slots);
    *out = s.release();
  });
}

int bm_server_submit(bm_server* server, const uint8_t* bone, const uint8_t* roi, uint64_t width,
                     uint64_t height, uint64_t depth, const double voxel_spacing[3],
                     int feature_mode, uint64_t* ticket) {
  return guarded([&] {
    if (!server || !server->impl || !ticket || !voxel_spacing) throw Error("null argument");
    *ticket = server->impl->submit(bone, roi, width, height, depth, voxel_spacing, feature_mode);
  });
}

int bm_server_poll(bm_server* server, uint64_t ticket, int wait, bm_batch_result* result,
                   bm_request_timing* timing, int* done) {
  return guarded([&] {
    if (!server || !server->impl || !done) throw Error("null argument");
    VolumeResult v;
    *done = server->impl->poll(ticket, wait != 0, &v, timing) ? 1 : 0;
    if (*done && result) to_c(v, result);
  });
}

int bm_server_cancel(bm_server* server, uint64_t ticket, int* cancelled) {
  return guarded([&] {
    if (!server || !server->impl) throw Error("null argument");
    const bool c = server->impl->cancel(ticket);
    if (cancelled) *cancelled = c ? 1 : 0;
  });
}

int bm_server_stats_get(bm_server* server, bm_server_stats* out) {
  return guarded([&] {
    if (!server || !server->impl || !out) throw Error("null argument");
    server->impl->stats(out);
  });
}

int bm_server_destroy(bm_server* server) {
  return guarded([&] {
    if (!server) return;
    delete server->impl;
    delete server;
  });
}

}  // extern "C"
