// bm_smooth.h -- the signed distance field and its Gaussian smoothing, the
// EDT consumers of the curvature and Minkowski measures of bone-morphometry
// 0.7.0-rc.3:
//   require_isotropic_spacing         src/sdf_curvature.rs:147-160
//   signed_distance_field_um          src/sdf_curvature.rs:162-180
//   gaussian_kernel                   src/sdf_curvature.rs:182-198
//   gaussian_smooth_nearest_in_place  src/sdf_curvature.rs:200-269
// and their composition as the three callers use it
// (src/sdf_curvature.rs:98-105, src/minkowski.rs:64-71, src/sdf_curvature.rs:719-726).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "bm_engine.h"

namespace bm {

// Largest Gaussian radius (taps 2r+1) accepted; beyond it std::range_error.
constexpr int64_t kMaxGaussRadius = 1 << 20;

// The isotropic spacing h, or Error with the reference's message.
double require_isotropic_spacing(const double spacing[3]);

// Normalised weights exp(-x^2 / 2 sigma^2), x = -r..r, r = floor(4 sigma + 0.5);
// {1} for sigma <= 0 (and for NaN, as Rust's saturating cast gives r = 0).
std::vector<double> gaussian_kernel(double sigma);

// All calls below run on stream s and return when their output is complete.
// They hold the engine's mutex (the caller takes it) and use its scratch slots.

// phi = +sqrt(d2_bone) h on bone voxels, -sqrt(d2_background) h elsewhere.
void signed_distance_field_device(Engine& e, const uint8_t* d_bone, const Dims& dims, double h,
                                  double* d_out, hipStream_t s);

// In-place x, y, z smoothing with clamped indices; d_data holds dims.n() values.
void gaussian_smooth_device(Engine& e, double* d_data, const Dims& dims, double sigma_vox,
                            hipStream_t s);

// signed_distance_field_device followed by gaussian_smooth_device, without
// writing the unsmoothed field: the x pass computes it from the two squared
// EDTs as it stages its lines.  With d_out null the result stays in the
// engine's scratch slot 2; returns where it is.
double* smoothed_signed_distance_field_device(Engine& e, const uint8_t* d_bone, const Dims& dims,
                                              double h, double sigma_vox, double* d_out,
                                              hipStream_t s);

// Dimensions accepted by the smoothing, which has no lower limit (a volume
// with no voxels is left as it is) and indexes with 64 bits.
Dims smooth_dims(uint64_t w, uint64_t h, uint64_t d);

}  // namespace bm
