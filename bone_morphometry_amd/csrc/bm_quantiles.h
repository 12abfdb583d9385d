// bm_quantiles.h -- exact order statistics of a thickness map on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm_engine.h"

namespace bm {

// The five linear quantiles of nonzero_stats (src/metrics.rs:242-255,271-291)
// over the positive finite values of map[n] (count of them known), each
// scaled as (value as f32 * scale_f32) as f64.
struct QuantileResult {
  double median = 0, p10 = 0, p25 = 0, p75 = 0, p90 = 0;
};

// top16 (device, may be null): counts of the values' top 16 bits over
// [kTopBase, kTopBase + kTopBins), from the clean-up kernel.
QuantileResult device_quantiles(Engine& e, const float* d_map, uint64_t n, uint64_t count,
                                float scale_f32, hipStream_t s, const uint32_t* top16 = nullptr);

}  // namespace bm
