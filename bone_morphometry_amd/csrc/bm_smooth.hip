// bm_smooth.hip -- the signed distance field of a bone volume and its
// separable Gaussian smoothing, in f64 (bone-morphometry 0.7.0-rc.3):
//   require_isotropic_spacing         src/sdf_curvature.rs:147-160
//   signed_distance_field_um          src/sdf_curvature.rs:162-180
//   gaussian_kernel                   src/sdf_curvature.rs:182-198
//   gaussian_smooth_nearest_in_place  src/sdf_curvature.rs:200-269
//
// The field is +sqrt(d2) h on bone voxels (d2 the squared EDT of the bone
// phase) and -sqrt(d2') h elsewhere (d2' that of the background phase).  Both
// EDTs run as the two items of one plan, in the same launches.
//
// The smoothing is three passes, x then y then z.  Each output is
//   acc = 0; for k in 0..2r: acc += w[k] * in[clamp(p + k - r)]
// with every product and sum rounded on its own (this file is built with
// contraction off, and the kernels say so again), so each output is the
// reference's to the bit.  A pass stages a tile of 64 lines x (64 + 2r)
// positions in LDS; each lane then streams its line through a register block
// of 8 outputs, so a staged value is read from LDS once per 8 outputs rather
// than once per tap.  Lines run across the 64 lanes in the y and z passes
// (x is contiguous in memory); in the x pass the tile is transposed on the
// way into LDS and back out, so global loads and stores stay contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "bm_device.h"
#include "bm_engine.h"
#include "bm_smooth.h"

namespace bm {

constexpr uint32_t kSmLanes = 64;   // lines per tile
constexpr uint32_t kSmTile = 64;    // outputs per line and tile
constexpr uint32_t kSmRegs = 8;     // outputs per lane and register block
constexpr uint32_t kSmWaves = 4;    // kSmTile / (kSmWaves * kSmRegs) = 2 blocks per wave
constexpr uint32_t kSmPitch = 65;   // LDS doubles per staged position (odd: no bank conflicts)

// Where a pass reads its input from.
struct SmoothSrc {
  const double* data = nullptr;      // f64 field
  const uint8_t* bone = nullptr;     // signed distance from the two squared EDTs:
  const uint32_t* d2_bone = nullptr;
  const uint32_t* d2_back = nullptr;
  double h = 0.0;
};

enum SrcKind { kSrcField = 0, kSrcSdf = 1 };

// src/sdf_curvature.rs:162-180: `(d2 as f64).sqrt() * h` on bone voxels and
// `-(d2 as f64).sqrt() * h` elsewhere, d2 the transform of the voxel's own
// phase; the other phase's transform is 0 there, so the sum is that d2 and
// no load depends on the bone byte.  The square root is correctly rounded,
// and -(a * h) == (-a) * h exactly.
__device__ __forceinline__ double sdf_value(uint8_t bone, uint32_t d2_bone, uint32_t d2_back,
                                            double h) {
  const double s = sqrt((double)(d2_bone + d2_back)) * h;
  return bone ? s : -s;
}

template <int kKind>
__device__ __forceinline__ double src_at(const SmoothSrc& s, uint64_t i) {
  if (kKind == kSrcField) return gbl(s.data)[i];
  return sdf_value(gbl(s.bone)[i], gbl(s.d2_bone)[i], gbl(s.d2_back)[i], s.h);
}

// A staged input held in registers between its loads and its use, so a lane
// issues all of a batch's loads before it waits for the first.
template <int kKind>
struct Staged {
  double v;
  __device__ __forceinline__ void load(const SmoothSrc& s, uint64_t i) { v = gbl(s.data)[i]; }
  __device__ __forceinline__ double value(const SmoothSrc&) const { return v; }
};
template <>
struct Staged<kSrcSdf> {
  uint32_t bone, d2_bone, d2_back;
  __device__ __forceinline__ void load(const SmoothSrc& s, uint64_t i) {
    bone = gbl(s.bone)[i];
    d2_bone = gbl(s.d2_bone)[i];
    d2_back = gbl(s.d2_back)[i];
  }
  __device__ __forceinline__ double value(const SmoothSrc& s) const {
    return sdf_value((uint8_t)bone, d2_bone, d2_back, s.h);
  }
};

__device__ __forceinline__ uint32_t clamp_line(int64_t v, uint32_t n) {
  return v < 0 ? 0u : (v >= (int64_t)n ? n - 1 : (uint32_t)v);
}

// Element index of (position a along the pass axis, line p, other o).
// Axis 0: a = x, p = y, o = z.  Axis 1: a = y, p = x, o = z.  Axis 2: a = z, p = x, o = y.
template <int kAxis>
__device__ __forceinline__ uint64_t smooth_index(uint32_t a, uint32_t p, uint32_t o, uint32_t W,
                                                 uint32_t H) {
  if (kAxis == 0) return ((uint64_t)o * H + p) * W + a;
  if (kAxis == 1) return ((uint64_t)o * H + a) * W + p;
  return ((uint64_t)a * H + o) * W + p;
}

// One register block: outputs o0..o0+7 of the lane's line, in reference
// order.  col[k * kSmPitch] is staged position o0 + k.  For output q the taps
// arrive as k = q + kk, kk = 0..K-1, in increasing k, so each output adds
// its terms in tap order.
__device__ __forceinline__ void smooth_block(const double* __restrict__ col,
                                             const double* __restrict__ w, uint32_t K,
                                             double (&acc)[kSmRegs]) {
#pragma clang fp contract(off)
#pragma unroll
  for (uint32_t q = 0; q < kSmRegs; ++q) acc[q] = 0.0;
  if (K >= kSmRegs) {
#pragma unroll
    for (uint32_t k = 0; k + 1 < kSmRegs; ++k) {  // head: outputs 0..k have begun
      const double v = col[k * kSmPitch];
#pragma unroll
      for (uint32_t q = 0; q <= k; ++q) acc[q] += w[k - q] * v;
    }
    double wv[kSmRegs];  // wv[q] = w[k - q]
#pragma unroll
    for (uint32_t q = 0; q < kSmRegs; ++q) wv[q] = w[kSmRegs - 1 - q];
    for (uint32_t k = kSmRegs - 1; k < K; ++k) {  // body: every output takes a tap
      const double v = col[k * kSmPitch];
#pragma unroll
      for (uint32_t q = 0; q < kSmRegs; ++q) acc[q] += wv[q] * v;
#pragma unroll
      for (uint32_t q = kSmRegs - 1; q > 0; --q) wv[q] = wv[q - 1];
      wv[0] = k + 1 < K ? w[k + 1] : 0.0;
    }
#pragma unroll
    for (uint32_t t = 0; t + 1 < kSmRegs; ++t) {  // tail: outputs t+1..7 still take taps
      const double v = col[(K + t) * kSmPitch];
#pragma unroll
      for (uint32_t q = t + 1; q < kSmRegs; ++q) acc[q] += w[K + t - q] * v;
    }
  } else {
#pragma unroll
    for (uint32_t q = 0; q < kSmRegs; ++q)
      for (uint32_t kk = 0; kk < K; ++kk) acc[q] += w[kk] * col[(q + kk) * kSmPitch];
  }
}

// One pass along kAxis.  A tile is 64 lines (p0..p0+63) x 64 outputs
// (a0..a0+63) at one value of the third coordinate; the launch covers
// ntiles = nA * nP * nO tiles, the tile along the axis fastest, spread over
// the XCDs in contiguous runs so the halos of neighbouring tiles meet in L2.
template <int kAxis, int kKind>
__global__ void __launch_bounds__(256)
k_smooth_tiled(SmoothSrc src, double* __restrict__ out, const double* __restrict__ w, uint32_t r,
               uint32_t W, uint32_t H, uint32_t D, uint32_t nA, uint32_t nP, uint32_t ntiles) {
#pragma clang fp contract(off)
  extern __shared__ double s_tile[];
  const uint32_t t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const uint32_t at = t % nA, rest = t / nA;
  const uint32_t pt = rest % nP, o = rest / nP;
  const uint32_t An = kAxis == 0 ? W : (kAxis == 1 ? H : D);
  const uint32_t Pn = kAxis == 0 ? H : W;
  const uint32_t a0 = at * kSmTile, p0 = pt * kSmLanes;
  const uint32_t K = 2 * r + 1, L = kSmTile + 2 * r;
  // the wave index is uniform; saying so keeps row tests scalar branches
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  BM_GLOBAL double* __restrict__ dst = gbl(out);

  // Stage positions a0 - r .. a0 + 63 + r (clamped) of the 64 lines: kStage
  // loads per lane in flight, then their values into LDS.  Nothing here is
  // conditional, so the compiler never waits for one batch before the next:
  // lines beyond the volume load its last line (their outputs are never
  // stored), and positions beyond the tile (j >= L) load a clamped element
  // and go to the spare row L.
  constexpr uint32_t kStage = 8;
  static_assert(kSmLanes % (kSmWaves * kStage) == 0, "rows per staging batch");
  if (kAxis == 0) {  // lines are rows: a wave takes whole rows, positions across
                     // lanes, so a row's base index is uniform and a lane only
                     // adds its clamped position
    for (uint32_t jb = 0; jb < L; jb += 64) {
      const uint32_t j = jb + lane;
      const uint32_t a = clamp_line((int64_t)a0 - r + j, An);
      const uint32_t row = j < L ? j : L;
      for (uint32_t pb = wave; pb < kSmLanes; pb += kSmWaves * kStage) {
        Staged<kKind> v[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u) {
          const uint32_t pl = min(p0 + pb + u * kSmWaves, Pn - 1);
          v[u].load(src, ((uint64_t)o * H + pl) * W + a);
        }
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u)
          s_tile[row * kSmPitch + pb + u * kSmWaves] = v[u].value(src);
      }
    }
  } else {  // lines are x-columns: one line per lane, positions uniform per wave
    const uint32_t pl = min(p0 + lane, Pn - 1);
    for (uint32_t jb = wave; jb < L; jb += kSmWaves * kStage) {
      Staged<kKind> v[kStage];
#pragma unroll
      for (uint32_t u = 0; u < kStage; ++u) {
        const uint32_t j = jb + u * kSmWaves;
        v[u].load(src, smooth_index<kAxis>(clamp_line((int64_t)a0 - r + j, An), pl, o, W, H));
      }
#pragma unroll
      for (uint32_t u = 0; u < kStage; ++u) {
        const uint32_t j = jb + u * kSmWaves;
        s_tile[(j < L ? j : L) * kSmPitch + lane] = v[u].value(src);
      }
    }
  }
  __syncthreads();

  constexpr uint32_t kBlocks = kSmTile / (kSmWaves * kSmRegs);
  double acc[kBlocks][kSmRegs];
#pragma unroll
  for (uint32_t b = 0; b < kBlocks; ++b) {
    const uint32_t o0 = (wave + b * kSmWaves) * kSmRegs;
    if (a0 + o0 < An) {
      smooth_block(s_tile + o0 * kSmPitch + lane, w, K, acc[b]);
    } else {
#pragma unroll
      for (uint32_t q = 0; q < kSmRegs; ++q) acc[b][q] = 0.0;
    }
  }

  if (kAxis == 0) {  // back through LDS, so rows are stored contiguously
    __syncthreads();
#pragma unroll
    for (uint32_t b = 0; b < kBlocks; ++b) {
      const uint32_t o0 = (wave + b * kSmWaves) * kSmRegs;
#pragma unroll
      for (uint32_t q = 0; q < kSmRegs; ++q) s_tile[(o0 + q) * kSmPitch + lane] = acc[b][q];
    }
    __syncthreads();
    const uint32_t a = a0 + lane;
    if (a < An)
      for (uint32_t p = wave; p < kSmLanes && p0 + p < Pn; p += kSmWaves)
        dst[smooth_index<0>(a, p0 + p, o, W, H)] = s_tile[lane * kSmPitch + p];
  } else {
    if (p0 + lane < Pn) {
#pragma unroll
      for (uint32_t b = 0; b < kBlocks; ++b) {
        const uint32_t o0 = (wave + b * kSmWaves) * kSmRegs;
#pragma unroll
        for (uint32_t q = 0; q < kSmRegs; ++q)
          if (a0 + o0 + q < An) dst[smooth_index<kAxis>(a0 + o0 + q, p0 + lane, o, W, H)] = acc[b][q];
      }
    }
  }
}

// A pass whose tile does not fit in LDS (radius beyond ~120): one output per
// thread, the taps read from memory.
template <int kKind>
__global__ void __launch_bounds__(256)
k_smooth_direct(SmoothSrc src, double* __restrict__ out, const double* __restrict__ w, uint32_t r,
                uint32_t W, uint32_t H, uint32_t D, int axis) {
#pragma clang fp contract(off)
  const uint64_t n = (uint64_t)W * H * D;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t K = 2 * r + 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t x = (uint32_t)(i % W);
    const uint64_t yz = i / W;
    const uint32_t y = (uint32_t)(yz % H), z = (uint32_t)(yz / H);
    const uint32_t a = axis == 0 ? x : (axis == 1 ? y : z);
    const uint32_t An = axis == 0 ? W : (axis == 1 ? H : D);
    const uint64_t step = axis == 0 ? 1 : (axis == 1 ? W : (uint64_t)W * H);
    const uint64_t base = i - (uint64_t)a * step;
    double acc = 0.0;
    for (uint32_t kk = 0; kk < K; ++kk) {
      const uint32_t b = clamp_line((int64_t)a + kk - r, An);
      acc += w[kk] * src_at<kKind>(src, base + (uint64_t)b * step);
    }
    gbl(out)[i] = acc;
  }
}

// The unsmoothed field, voxel by voxel.
__global__ void __launch_bounds__(256)
k_sdf(const uint8_t* __restrict__ bone, const uint32_t* __restrict__ d2_bone,
      const uint32_t* __restrict__ d2_back, uint64_t n, double h, double* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    gbl(out)[i] = sdf_value(gbl(bone)[i], gbl(d2_bone)[i], gbl(d2_back)[i], h);
}

// ---------------------------------------------------------------------------
// Host side.

// f64 as Rust's Display ({}) writes it: the shortest digits that read back,
// as a plain decimal without exponent.
static std::string rust_display_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  for (int prec = 0; prec <= 16; ++prec) {
    snprintf(buf, sizeof buf, "%.*e", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s = buf;
  const bool neg = s[0] == '-';
  if (neg) s = s.substr(1);
  const size_t e = s.find('e');
  const int exp10 = std::atoi(s.c_str() + e + 1);
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string r;
  const int point = exp10 + 1;  // digits before the decimal point
  if (point <= 0) {
    r = "0." + std::string(-point, '0') + digits;
  } else if ((size_t)point >= digits.size()) {
    r = digits + std::string(point - digits.size(), '0');
  } else {
    r = digits.substr(0, point) + "." + digits.substr(point);
  }
  if (r == "0") return neg ? "-0" : "0";
  return neg ? "-" + r : r;
}

double require_isotropic_spacing(const double spacing[3]) {
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(spacing[i]) || spacing[i] <= 0.0)
      throw Error("voxel spacing must be finite and positive");
  const double h = spacing[0];
  const double tol = 1e-9 * std::max(std::fabs(h), 1.0);
  if (std::fabs(spacing[1] - h) > tol || std::fabs(spacing[2] - h) > tol)
    throw Error("advanced SDF curvature currently requires isotropic spacing; got [" +
                rust_display_f64(spacing[0]) + ", " + rust_display_f64(spacing[1]) + ", " +
                rust_display_f64(spacing[2]) + "] \xC2\xB5m");
  return h;
}

std::vector<double> gaussian_kernel(double sigma) {
  if (sigma <= 0.0) return {1.0};
  const double rf = std::floor(4.0 * sigma + 0.5);
  int64_t radius = 0;  // `as isize`: NaN -> 0
  if (!std::isnan(rf)) {
    if (rf > (double)kMaxGaussRadius)
      throw std::range_error("gaussian kernel radius " + rust_display_f64(rf) +
                             " exceeds this build's limit of " + std::to_string(kMaxGaussRadius) +
                             " voxels");
    radius = (int64_t)rf;
  }
  std::vector<double> kernel;
  kernel.reserve((size_t)(2 * radius + 1));
  double sum = 0.0;
  for (int64_t i = -radius; i <= radius; ++i) {
    const double x = (double)i;
    const double w = std::exp(-0.5 * x * x / (sigma * sigma));
    kernel.push_back(w);
    sum += w;
  }
  for (double& w : kernel) w /= sum;
  return kernel;
}

// The staged tile (64 + 2r positions) and a spare row.
static size_t tiled_lds_bytes(uint32_t r) {
  return (size_t)(kSmTile + 2 * r + 1) * kSmPitch * sizeof(double);
}

template <int kAxis, int kKind>
static void launch_tiled(const SmoothSrc& src, double* out, const double* d_w, uint32_t r,
                         const Dims& dims, hipStream_t s) {
  const uint32_t An = kAxis == 0 ? dims.w : (kAxis == 1 ? dims.h : dims.d);
  const uint32_t Pn = kAxis == 0 ? dims.h : dims.w;
  const uint32_t On = kAxis == 2 ? dims.h : dims.d;
  const uint32_t nA = (An + kSmTile - 1) / kSmTile, nP = (Pn + kSmLanes - 1) / kSmLanes;
  const uint64_t ntiles = (uint64_t)nA * nP * On;
  if (ntiles >= (1ull << 31)) throw std::range_error("smoothing pass exceeds the launch limit");
  const uint32_t per = (uint32_t)((ntiles + kXcds - 1) / kXcds);
  const char* names[3] = {"smooth_x", "smooth_y", "smooth_z"};
  BM_LAUNCH(names[kAxis], (k_smooth_tiled<kAxis, kKind>), per * kXcds, 256, tiled_lds_bytes(r),
            s, src, out, d_w, r, dims.w, dims.h, dims.d, nA, nP, (uint32_t)ntiles);
  BM_HIP(hipGetLastError());
}

// LDS a block may allocate on the current device (160 KiB on gfx950).
static size_t lds_per_block() {
  int dev = 0, v = 0;
  BM_HIP(hipGetDevice(&dev));
  BM_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  return (size_t)v;
}

template <int kKind>
static void smooth_pass(int axis, const SmoothSrc& src, double* out, const double* d_w,
                        uint32_t r, const Dims& dims, hipStream_t s) {
  if (tiled_lds_bytes(r) <= lds_per_block()) {
    if (axis == 0) launch_tiled<0, kKind>(src, out, d_w, r, dims, s);
    else if (axis == 1) launch_tiled<1, kKind>(src, out, d_w, r, dims, s);
    else launch_tiled<2, kKind>(src, out, d_w, r, dims, s);
    return;
  }
  const uint64_t n = dims.n();
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
  BM_LAUNCH("smooth_direct", k_smooth_direct<kKind>, blocks, 256, 0, s, src, out, d_w, r, dims.w,
            dims.h, dims.d, axis);
  BM_HIP(hipGetLastError());
}

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// The weights in device memory (the engine's table slot); returns null when
// the kernel is the identity.
static const double* upload_kernel(Engine& e, const std::vector<double>& k, hipStream_t s) {
  if (k.size() <= 1) return nullptr;
  double* d_w = static_cast<double*>(e.scratch_bytes(k.size() * sizeof(double), 3));
  BM_HIP(hipMemcpyAsync(d_w, k.data(), k.size() * sizeof(double), hipMemcpyHostToDevice, s));
  return d_w;
}

// Both squared EDTs into one buffer of 2n u32 (bone phase first), in one
// two-item plan.  The foreground's error is the one reported first, as the
// reference computes that transform first.
static void both_edts(Engine& e, const uint8_t* d_bone, const Dims& dims, uint32_t* d2,
                      hipStream_t s) {
  ItemSpec f, b;
  f.src.data = b.src.data = d_bone;
  f.src.want_nonzero = 1;
  b.src.want_nonzero = 0;
  f.dims = b.dims = dims;
  f.tune = b.tune = tuned_for(dims);
  f.thickness = b.thickness = false;
  f.d2_out = d2;
  b.d2_out = d2 + dims.n();
  Plan& pl = e.plain_plan();
  pl.build({f, b}, s);
  pl.launch(s);
  pl.fetch(s);
  for (int i = 0; i < 2; ++i)
    if (!pl.result(i).status.has_site)
      throw Error(std::string("distance transform is undefined for an all-") +
                  (i == 0 ? "foreground" : "background") +
                  " volume unless an exterior convention is specified");
}

void signed_distance_field_device(Engine& e, const uint8_t* d_bone, const Dims& dims, double h,
                                  double* d_out, hipStream_t s) {
  const uint64_t n = dims.n();
  uint32_t* d2 = static_cast<uint32_t*>(e.scratch_bytes(n * 8, 1));
  both_edts(e, d_bone, dims, d2, s);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
  BM_LAUNCH("sdf", k_sdf, blocks, 256, 0, s, d_bone, d2, d2 + n, n, h, d_out);
  BM_HIP(hipGetLastError());
  BM_HIP(hipStreamSynchronize(s));
}

void gaussian_smooth_device(Engine& e, double* d_data, const Dims& dims, double sigma_vox,
                            hipStream_t s) {
  const std::vector<double> k = gaussian_kernel(sigma_vox);
  if (k.size() <= 1 || dims.n() == 0) return;
  const uint32_t r = (uint32_t)(k.size() / 2);
  const uint64_t n = dims.n();
  double* t1 = static_cast<double*>(e.scratch_bytes(n * 8, 1));
  double* t2 = static_cast<double*>(e.scratch_bytes(n * 8, 2));
  const double* d_w = upload_kernel(e, k, s);
  SmoothSrc a, b, c;
  a.data = d_data;
  b.data = t1;
  c.data = t2;
  smooth_pass<kSrcField>(0, a, t1, d_w, r, dims, s);
  smooth_pass<kSrcField>(1, b, t2, d_w, r, dims, s);
  smooth_pass<kSrcField>(2, c, d_data, d_w, r, dims, s);
  BM_HIP(hipStreamSynchronize(s));
}

double* smoothed_signed_distance_field_device(Engine& e, const uint8_t* d_bone, const Dims& dims,
                                              double h, double sigma_vox, double* d_out,
                                              hipStream_t s) {
  const std::vector<double> k = gaussian_kernel(sigma_vox);
  const uint64_t n = dims.n();
  // slot 1: the two squared EDTs (2n u32), later the y pass's output (n f64);
  // slot 2: the x pass's output, later the result when d_out is null
  uint32_t* d2 = static_cast<uint32_t*>(e.scratch_bytes(n * 8, 1));
  double* t2 = static_cast<double*>(e.scratch_bytes(n * 8, 2));
  if (!d_out) d_out = t2;
  both_edts(e, d_bone, dims, d2, s);
  if (k.size() <= 1) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    BM_LAUNCH("sdf", k_sdf, blocks, 256, 0, s, d_bone, d2, d2 + n, n, h, d_out);
    BM_HIP(hipGetLastError());
    BM_HIP(hipStreamSynchronize(s));
    return d_out;
  }
  const uint32_t r = (uint32_t)(k.size() / 2);
  const double* d_w = upload_kernel(e, k, s);
  SmoothSrc a, b, c;
  a.bone = d_bone;
  a.d2_bone = d2;
  a.d2_back = d2 + n;
  a.h = h;
  b.data = t2;
  c.data = reinterpret_cast<double*>(d2);
  smooth_pass<kSrcSdf>(0, a, t2, d_w, r, dims, s);
  smooth_pass<kSrcField>(1, b, reinterpret_cast<double*>(d2), d_w, r, dims, s);
  smooth_pass<kSrcField>(2, c, d_out, d_w, r, dims, s);
  BM_HIP(hipStreamSynchronize(s));
  return d_out;
}

Dims smooth_dims(uint64_t w, uint64_t h, uint64_t d) {
  if (w > 0x7FFFFFFFull || h > 0x7FFFFFFFull || d > 0x7FFFFFFFull ||
      (unsigned __int128)w * h * d > (1ull << 36))
    throw std::range_error("volume " + std::to_string(w) + "x" + std::to_string(h) + "x" +
                           std::to_string(d) + " exceeds the smoothing's limits (2^36 voxels)");
  Dims dims;
  dims.w = (uint32_t)w;
  dims.h = (uint32_t)h;
  dims.d = (uint32_t)d;
  return dims;
}

}  // namespace bm
