// bm_analyze.h -- the analyze() pieces shared by the one-volume calls
// (bm_api.cpp) and the batched and serving forms (bm_serve.cpp).
#pragma once

#include <string>
#include <vector>

#include "../../include/bm.h"
#include "bm_engine.h"

namespace bm {

// Dimension checks of BinaryVolume::new (src/volume.rs:27-34) plus the device
// path's limits (std::range_error beyond them).
Dims device_dims(uint64_t w, uint64_t h, uint64_t d);
// Spacing check of analyze (src/metrics.rs:71-78) and the feature mode range.
void check_spacing(const double spacing[3]);
void check_mode(int mode);
// The bone and marrow items of one analyze call on device-resident inputs.
// keep_maps: both thickness maps stay in the plan (for sequential_sums).
void analyze_items(const uint8_t* d_bone, const uint8_t* d_roi, const Dims& dims,
                   const double spacing[3], int mode, std::vector<ItemSpec>& out,
                   bool keep_maps = false);
// analyze's result from its two items, with the error precedence of
// src/metrics.rs:90-104 (throws Error with the reference's message).  With
// seq (two entries, bone then marrow), mean and SD come from those sums
// instead of the items' exact device sums.
void analyze_finish(const ItemResult& bone, const ItemResult& marrow, uint64_t n,
                    const double spacing[3], int mode, bm_thickness_spacing* out,
                    const StatsResult* seq = nullptr);
// The order of the f64 sums behind mean and SD (BM_STATS_* of bm.h).
bool stats_sequential();
// nonzero_stats' count, sums and maximum (src/metrics.rs:215-226) over a
// device map, in the reference's order: linear index order, sequential f64.
StatsResult sequential_sums(const float* d_map, uint64_t n, float scale_f32, hipStream_t s);
// Both items' sequential sums (bone, marrow) of a plan built with keep_maps,
// items first and first + 1; false (nothing computed) when a phase is empty,
// which analyze_finish reports.
bool replay_sums(const Plan& plan, size_t first, float scale_f32, hipStream_t s,
                 StatsResult out[2]);
// The autotuner (bm_tune.cpp): one sample per shape class, device-resident.
struct TuneSample {
  const uint8_t* bone;
  const uint8_t* roi;
  Dims dims;
};
void autotune(const std::vector<TuneSample>& samples, const std::string& path, bool verbose);
void tuning_load(const std::string& path);

// Runs f and maps an exception to a BM_ERR_* code and message (BM_OK, "" if
// none was thrown).
int capture_errors(void (*f)(void*), void* ctx, std::string* message);

}  // namespace bm
