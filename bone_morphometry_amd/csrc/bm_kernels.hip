// bm_kernels.hip -- gfx950 kernels of the exact squared EDT and local thickness.
//
// EDT (src/edt.rs:57-156).  The reference runs an f64 Felzenszwalb-Huttenlocher
// lower envelope per line; on binary input its result is the exact integer
// squared distance to the nearest non-selected voxel, so any exact integer
// method is bit-identical.  Here:
//   X pass  one wavefront per row: ballots give each voxel its nearest site to
//           the left and right; output is the 1-D distance as u16.
//   Y pass  64 x-columns of a z-slice staged in LDS; each lane scans offsets
//           o = 1, 2, ... around its row and stops once o^2 >= best, which is
//           exact because every candidate at offset >= o adds at least o^2.
//   Z pass  the same search along z over 32 x-columns; writes the squared EDT
//           (0 for non-selected voxels) and the maximum.
// Local thickness (src/local_thickness.rs:16-44): templates and diameters as
// look-up tables indexed by r^2, the 26-neighbour ridge filter, the sphere
// fill, the surface clean-up with the phase mask, and exact statistics.
#include <hip/hip_runtime.h>

#include "bm_engine.h"

namespace bm {

typedef unsigned __int128 u128;

__device__ __forceinline__ uint32_t sq16(uint32_t d) { return d == kInf16 ? kInf32 : d * d; }

__device__ __forceinline__ bool selected_voxel(uint8_t v, uint8_t r, uint32_t want) {
  return r != 0 && ((v != 0) == (want != 0));
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}

// ---------------------------------------------------------------------------
// X pass.  Sites bits of 8 consecutive voxels x..x+7 of a row; bit j set when
// voxel x+j exists and is not selected.  Counts selected voxels.
__device__ __forceinline__ uint32_t site_bits8(const uint8_t* __restrict__ data,
                                               const uint8_t* __restrict__ roi, uint32_t want,
                                               size_t rb, uint32_t x, uint32_t W,
                                               uint32_t& selcnt) {
  uint32_t bits = 0;
  const bool aligned = ((reinterpret_cast<uintptr_t>(data + rb + x) |
                         (roi ? reinterpret_cast<uintptr_t>(roi + rb + x) : 0)) & 7) == 0;
  if (x + 8 <= W && aligned) {
    uint2 v = *reinterpret_cast<const uint2*>(data + rb + x);
    uint2 r = roi ? *reinterpret_cast<const uint2*>(roi + rb + x) : make_uint2(0x01010101u, 0x01010101u);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      uint8_t vb = (uint8_t)((j < 4 ? v.x : v.y) >> (8 * (j & 3)));
      uint8_t rbt = (uint8_t)((j < 4 ? r.x : r.y) >> (8 * (j & 3)));
      if (selected_voxel(vb, rbt, want)) ++selcnt;
      else bits |= 1u << j;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (x + j < W) {
        uint8_t vb = data[rb + x + j];
        uint8_t rbt = roi ? roi[rb + x + j] : (uint8_t)1;
        if (selected_voxel(vb, rbt, want)) ++selcnt;
        else bits |= 1u << j;
      }
    }
  }
  return bits;
}

constexpr int kNoneL = -(1 << 20);
constexpr int kNoneR = (1 << 21);

__global__ void __launch_bounds__(256)
k_edt_x(const uint8_t* __restrict__ data, const uint8_t* __restrict__ roi, uint32_t want,
        uint16_t* __restrict__ out, uint32_t W, uint32_t rows, PhaseStatus* __restrict__ st) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nseg = (W + 511) / 512;
  uint32_t selcnt = 0, anysite = 0;
  for (uint32_t r = gw; r < rows; r += nw) {
    const size_t rb = (size_t)r * W;
    int carryL = kNoneL;
    for (uint32_t s = 0; s < nseg; ++s) {
      const uint32_t x0 = s * 512 + lane * 8;
      const uint32_t bits = site_bits8(data, roi, want, rb, x0, W, selcnt);
      anysite |= bits;
      const unsigned long long has = __ballot(bits != 0);
      const int last = bits ? (int)(x0 + 31 - __clz(bits)) : kNoneL;
      const int first = bits ? (int)(x0 + __ffs(bits) - 1) : kNoneR;
      const unsigned long long lower = has & ((1ull << lane) - 1ull);
      const int fromL = __shfl(last, lower ? 63 - __clzll(lower) : 0);
      const int Lc = lower ? fromL : carryL;
      const unsigned long long higher = lane == 63 ? 0ull : (has & (~0ull << (lane + 1)));
      const int fromR = __shfl(first, higher ? __ffsll(higher) - 1 : 0);
      const int Rc = higher ? fromR : kNoneR;
      if (has) carryL = __shfl(last, 63 - __clzll(has));
      if (x0 < W) {
        uint16_t o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int x = (int)x0 + j;
          const uint32_t mL = bits & ((2u << j) - 1u);
          const int L = mL ? (int)(x0 + 31 - __clz(mL)) : Lc;
          const uint32_t mR = bits & (~0u << j);
          const int R = mR ? (int)(x0 + __ffs(mR) - 1) : Rc;
          const int dist = min(x - L, R - x);
          o[j] = dist > 65534 ? (uint16_t)kInf16 : (uint16_t)dist;
        }
        uint16_t* p = out + rb + x0;
        if (x0 + 8 <= W && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
          uint4 v;
          v.x = o[0] | ((uint32_t)o[1] << 16);
          v.y = o[2] | ((uint32_t)o[3] << 16);
          v.z = o[4] | ((uint32_t)o[5] << 16);
          v.w = o[6] | ((uint32_t)o[7] << 16);
          *reinterpret_cast<uint4*>(p) = v;
        } else {
          for (int j = 0; j < 8; ++j)
            if (x0 + j < W) p[j] = o[j];
        }
      }
    }
    if (nseg > 1) {
      // Sites right of each segment: second sweep from the right end.
      int carryR = kNoneR;
      for (uint32_t s = nseg; s-- > 0;) {
        const uint32_t x0 = s * 512 + lane * 8;
        uint32_t dummy = 0;
        const uint32_t bits = site_bits8(data, roi, want, rb, x0, W, dummy);
        const unsigned long long has = __ballot(bits != 0);
        const int first = bits ? (int)(x0 + __ffs(bits) - 1) : kNoneR;
        const unsigned long long higher = lane == 63 ? 0ull : (has & (~0ull << (lane + 1)));
        const int fromR = __shfl(first, higher ? __ffsll(higher) - 1 : 0);
        const int Rc = higher ? fromR : carryR;
        if (has) carryR = __shfl(first, __ffsll(has) - 1);
        if (x0 < W) {
          uint16_t* p = out + rb + x0;
          for (int j = 0; j < 8; ++j) {
            if (x0 + j >= W) break;
            const uint32_t mR = bits & (~0u << j);
            if (mR) continue;  // a site inside the segment is already accounted for
            const int dist = Rc - (int)(x0 + j);
            if (dist <= 65534 && (uint32_t)dist < p[j]) p[j] = (uint16_t)dist;
          }
        }
      }
    }
  }
  const unsigned long long tot = wave_sum_u64(selcnt);
  const unsigned long long anyb = __ballot(anysite != 0);
  if (lane == 0) {
    if (tot) atomicAdd(&st->selected, tot);
    if (anyb) atomicOr(&st->has_site, 1u);
  }
}

// ---------------------------------------------------------------------------
// Y pass: tile of 64 x-columns x rows [ya, yb) of one z-slice in LDS (u16).
__global__ void __launch_bounds__(256)
k_edt_y(const uint16_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t W, uint32_t H,
        uint32_t chunk, uint32_t halo) {
  extern __shared__ uint16_t ytile[];
  const uint32_t x0 = blockIdx.x * 64, z = blockIdx.y;
  const uint32_t y0 = blockIdx.z * chunk, y1 = min(H, y0 + chunk);
  const uint32_t ya = y0 > halo ? y0 - halo : 0, yb = min(H, y1 + halo);
  const size_t sbase = (size_t)z * W * H;
  const uint32_t nrows = yb - ya;
  const bool vec = (x0 + 64 <= W) && ((W & 7) == 0);
  for (uint32_t i = threadIdx.x; i < nrows * 8; i += blockDim.x) {
    const uint32_t rr = i >> 3, c = (i & 7) * 8;
    const uint16_t* p = in + sbase + (size_t)(ya + rr) * W + x0 + c;
    uint4 v;
    if (vec) {
      v = *reinterpret_cast<const uint4*>(p);
    } else {
      uint32_t t[8];
      for (int k = 0; k < 8; ++k) t[k] = (x0 + c + k < W) ? p[k] : kInf16;
      v = make_uint4(t[0] | (t[1] << 16), t[2] | (t[3] << 16), t[4] | (t[5] << 16),
                     t[6] | (t[7] << 16));
    }
    *reinterpret_cast<uint4*>(&ytile[rr * 64 + c]) = v;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const uint32_t x = x0 + lane;
  if (x >= W) return;
  const uint16_t* col = in + sbase + x;
  for (uint32_t q = y0 + wv; q < y1; q += nwv) {
    uint32_t best = sq16(ytile[(q - ya) * 64 + lane]);
    for (uint32_t o = 1; o * o < best; ++o) {
      bool any = false;
      if (q >= o) {
        const uint32_t p = q - o;
        const uint32_t f = p >= ya ? sq16(ytile[(p - ya) * 64 + lane]) : sq16(col[(size_t)p * W]);
        best = min(best, f + o * o);
        any = true;
      }
      if (q + o < H) {
        const uint32_t p = q + o;
        const uint32_t f = p < yb ? sq16(ytile[(p - ya) * 64 + lane]) : sq16(col[(size_t)p * W]);
        best = min(best, f + o * o);
        any = true;
      }
      if (!any) break;
    }
    out[sbase + (size_t)q * W + x] = best;
  }
}

// ---------------------------------------------------------------------------
// Z pass: tile of 32 x-columns x z-range [za, zb) at one y in LDS (u32).
__global__ void __launch_bounds__(256)
k_edt_z(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t W, uint32_t H,
        uint32_t D, uint32_t chunk, uint32_t halo, PhaseStatus* __restrict__ st) {
  extern __shared__ uint32_t ztile[];
  const uint32_t x0 = blockIdx.x * 32, y = blockIdx.y;
  const uint32_t z0 = blockIdx.z * chunk, z1 = min(D, z0 + chunk);
  const uint32_t za = z0 > halo ? z0 - halo : 0, zb = min(D, z1 + halo);
  const size_t slice = (size_t)W * H;
  const size_t base = (size_t)y * W + x0;
  const uint32_t nrows = zb - za;
  const bool vec = (x0 + 32 <= W) && ((W & 3) == 0);
  for (uint32_t i = threadIdx.x; i < nrows * 8; i += blockDim.x) {
    const uint32_t rr = i >> 3, c = (i & 7) * 4;
    const uint32_t* p = in + (size_t)(za + rr) * slice + base + c;
    uint4 v;
    if (vec) {
      v = *reinterpret_cast<const uint4*>(p);
    } else {
      v.x = (x0 + c + 0 < W) ? p[0] : kInf32;
      v.y = (x0 + c + 1 < W) ? p[1] : kInf32;
      v.z = (x0 + c + 2 < W) ? p[2] : kInf32;
      v.w = (x0 + c + 3 < W) ? p[3] : kInf32;
    }
    *reinterpret_cast<uint4*>(&ztile[rr * 32 + c]) = v;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const uint32_t xl = lane & 31, zs = lane >> 5;
  const uint32_t x = x0 + xl;
  uint32_t mx = 0;
  if (x < W) {
    const uint32_t* col = in + base + xl;
    for (uint32_t q = z0 + 2 * wv + zs; q < z1; q += 2 * nwv) {
      uint32_t best = ztile[(q - za) * 32 + xl];
      for (uint32_t o = 1; o * o < best; ++o) {
        bool any = false;
        if (q >= o) {
          const uint32_t p = q - o;
          const uint32_t f = p >= za ? ztile[(p - za) * 32 + xl] : col[(size_t)p * slice];
          best = min(best, f + o * o);
          any = true;
        }
        if (q + o < D) {
          const uint32_t p = q + o;
          const uint32_t f = p < zb ? ztile[(p - za) * 32 + xl] : col[(size_t)p * slice];
          best = min(best, f + o * o);
          any = true;
        }
        if (!any) break;
      }
      out[(size_t)q * slice + base + xl] = best;
      mx = max(mx, best);
    }
  }
  mx = wave_max_u32(mx);
  if (lane == 0 && mx) atomicMax(&st->max_d2, mx);
}

// ---------------------------------------------------------------------------
// integer_sqrt (src/local_thickness.rs:338-348): exact floor(sqrt(v)).
__device__ __forceinline__ uint32_t isqrt_u32(uint32_t v) {
  uint32_t r = (uint32_t)sqrtf((float)v);
  while ((uint64_t)(r + 1) * (r + 1) <= v) ++r;
  while ((uint64_t)r * r > v) --r;
  return r;
}
__device__ __forceinline__ uint32_t ceil_sqrt_u32(uint32_t v) {
  const uint32_t f = isqrt_u32(v);
  return f * f == v ? f : f + 1;
}

// scan_cube (src/local_thickness.rs:75-102) for one r^2.
__device__ uint32_t scan_cube(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t r2) {
  const uint32_t radius = isqrt_u32(r2) + 1;
  uint32_t maximum = 0;
  for (uint32_t k = 0; k <= radius; ++k) {
    const uint32_t sk = k * k, dk = (k + dz) * (k + dz);
    for (uint32_t j = 0; j <= radius; ++j) {
      const uint32_t skj = sk + j * j;
      if (skj > r2) break;  // j*j grows with j: the rest of the row is outside too
      const uint32_t ip = isqrt_u32(r2 - skj) + dx;
      maximum = max(maximum, dk + (j + dy) * (j + dy) + ip * ip);
    }
  }
  return maximum;
}

// Diameter (2 * sqrt_f64(m)) as f32 (src/local_thickness.rs:206).  For
// m < 2^24, (float)m is exact and the correctly rounded f32 sqrt scaled by 2
// equals RN32(2 sqrt(m)); the reference's RN32(RN64(2 sqrt(m))) is the same
// value because 2 sqrt(m) never lies within 2^-53 (relative) of an f32
// rounding midpoint.  Larger m take the f64 route.
__device__ __forceinline__ float diameter_of(uint32_t m) {
  if (m < (1u << 24)) return 2.0f * sqrtf((float)m);
  return (float)(2.0 * sqrt((double)m));
}

// Templates T_c[r^2] and diameters for every r^2 in [0, max_d2].
__global__ void __launch_bounds__(256)
k_luts(const PhaseStatus* __restrict__ st, uint32_t bound, uint32_t* __restrict__ tmpl,
       float* __restrict__ diam) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > bound || !st->has_site) return;
  const uint32_t mx = st->max_d2;
  if (m > mx) return;
  diam[m] = diameter_of(m);
  if (m == 0) return;
  tmpl[m] = scan_cube(1, 0, 0, m);
  tmpl[(bound + 1) + m] = scan_cube(1, 1, 0, m);
  tmpl[2 * (bound + 1) + m] = scan_cube(1, 1, 1, m);
}

// ---------------------------------------------------------------------------
// distance_ridges (src/local_thickness.rs:104-167): a selected voxel (r^2 > 0)
// is a ridge unless an in-bounds 26-neighbour has d2 >= T_c(r^2), c = number of
// nonzero offset components.
__global__ void __launch_bounds__(256)
k_ridges(const uint32_t* __restrict__ d2, uint32_t W, uint32_t H, uint32_t D,
         const uint32_t* __restrict__ tmpl, uint32_t bound, PhaseStatus* __restrict__ st,
         uint2* __restrict__ list) {
  if (!st->has_site) return;
  const uint64_t n = (uint64_t)W * H * D;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t idx = base + threadIdx.x;
    bool ridge = false;
    uint32_t r2 = 0;
    if (idx < n) {
      r2 = d2[idx];
      if (r2 != 0) {
        const uint32_t x = (uint32_t)(idx % W);
        const uint64_t t = idx / W;
        const uint32_t y = (uint32_t)(t % H), z = (uint32_t)(t / H);
        const uint32_t t1 = tmpl[r2], t2 = tmpl[(bound + 1) + r2], t3 = tmpl[2 * (bound + 1) + r2];
        ridge = true;
        for (int dz = -1; dz <= 1 && ridge; ++dz) {
          const int nz = (int)z + dz;
          if (nz < 0 || nz >= (int)D) continue;
          for (int dy = -1; dy <= 1 && ridge; ++dy) {
            const int ny = (int)y + dy;
            if (ny < 0 || ny >= (int)H) continue;
            for (int dx = -1; dx <= 1; ++dx) {
              if (dx == 0 && dy == 0 && dz == 0) continue;
              const int nx = (int)x + dx;
              if (nx < 0 || nx >= (int)W) continue;
              const int c = (dx != 0) + (dy != 0) + (dz != 0);
              const uint32_t tv = c == 1 ? t1 : (c == 2 ? t2 : t3);
              if (d2[((size_t)nz * H + ny) * W + nx] >= tv) {
                ridge = false;
                break;
              }
            }
          }
        }
      }
    }
    const unsigned long long m = __ballot(ridge);
    if (m) {
      unsigned long long slot0 = 0;
      if (lane == 0) slot0 = atomicAdd(&st->ridge_count, (unsigned long long)__popcll(m));
      slot0 = __shfl(slot0, 0);
      if (ridge) {
        const unsigned long long below = m & ((1ull << lane) - 1ull);
        list[slot0 + __popcll(below)] = make_uint2((uint32_t)idx, r2);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// propagate_ridges (src/local_thickness.rs:169-203) as an atomic-max scatter:
// one wavefront per ridge, lanes over the (2R+1)^2 (y, x) plane of each z.
__global__ void __launch_bounds__(256)
k_fill_scatter(const uint2* __restrict__ list, const PhaseStatus* __restrict__ st, uint32_t W,
               uint32_t H, uint32_t D, uint32_t* __restrict__ M) {
  if (!st->has_site) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t count = st->ridge_count;
  for (uint64_t r = gw; r < count; r += nw) {
    const uint2 e = list[r];
    const uint32_t r2 = e.y;
    const uint32_t x = e.x % W;
    const uint32_t t = e.x / W;
    const uint32_t y = t % H, z = t / H;
    const int R = (int)ceil_sqrt_u32(r2);
    const int side = 2 * R + 1;
    const int plane = side * side;
    for (int dz = -R; dz <= R; ++dz) {
      const int nz = (int)z + dz;
      const uint32_t dz2 = (uint32_t)(dz * dz);
      if (nz < 0 || nz >= (int)D || dz2 > r2) continue;
      for (int k = lane; k < plane; k += 64) {
        const int dy = k / side - R, dx = k % side - R;
        const int ny = (int)y + dy, nx = (int)x + dx;
        if (ny < 0 || ny >= (int)H || nx < 0 || nx >= (int)W) continue;
        if (dz2 + (uint32_t)(dy * dy) + (uint32_t)(dx * dx) > r2) continue;
        atomicMax(&M[((size_t)nz * H + ny) * W + nx], r2);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// NEIGHBORS_26 in Fiji's visiting order (src/local_thickness.rs:283-312); the
// f32 sum of the clean-up runs in this order.
__constant__ int8_t c_nb26[26][3] = {
    {0, 0, -1},  {0, 0, 1},   {0, -1, 0},  {0, 1, 0},   {-1, 0, 0},  {1, 0, 0},  {0, 1, -1},
    {0, 1, 1},   {1, -1, 0},  {1, 1, 0},   {-1, 0, 1},  {1, 0, 1},   {0, -1, -1}, {0, -1, 1},
    {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1}, {1, 0, -1},  {1, 1, 1},   {1, -1, 1},  {-1, 1, 1},
    {-1, -1, 1}, {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};

// Pass 1 of clean_up_surface_jaggies (src/local_thickness.rs:214-238): a
// nonzero voxel with an in-bounds 26-neighbour equal to zero is a surface voxel.
// Out-of-bounds reads are -1 (src/local_thickness.rs:325-327), never zero.
__device__ __forceinline__ bool surface_voxel(const uint32_t* __restrict__ M, int x, int y, int z,
                                              int W, int H, int D) {
  if (M[((size_t)z * H + y) * W + x] == 0) return false;
  for (int dz = -1; dz <= 1; ++dz) {
    const int nz = z + dz;
    if (nz < 0 || nz >= D) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = y + dy;
      if (ny < 0 || ny >= H) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int nx = x + dx;
        if ((dx | dy | dz) == 0 || nx < 0 || nx >= W) continue;
        if (M[((size_t)nz * H + ny) * W + nx] == 0) return true;
      }
    }
  }
  return false;
}

// Fixed-point accumulation of one value (positive finite f32).
__device__ __forceinline__ void acc_value(float v, int lsb_exp, unsigned long long& cnt, u128& s,
                                          u128& s2, uint32_t& mxb, uint32_t& flags) {
  const uint32_t bits = __float_as_uint(v);
  const int e = (int)((bits >> 23) & 0xFF);
  const uint32_t mant = (bits & 0x7FFFFFu) | (e ? 0x800000u : 0u);
  // Values lie in [2s, 2^17 s] for scale s, so shift is in [4, 21]; with
  // shift <= 24 each k^2 < 2^96 and 2^32 of them cannot overflow 128 bits.
  const int shift = (e ? e : 1) - 150 - lsb_exp;
  if (shift < 0 || shift > 24) {
    flags |= 1u;  // outside the exact fixed-point range
    return;
  }
  const u128 k = (u128)mant << shift;
  cnt += 1;
  s += k;
  s2 += (u128)((unsigned long long)mant * mant) << (2 * shift);
  mxb = max(mxb, bits);
}

__device__ __forceinline__ u128 shfl_xor_u128(u128 v, int off) {
  unsigned long long lo = (unsigned long long)v, hi = (unsigned long long)(v >> 64);
  lo = __shfl_xor(lo, off);
  hi = __shfl_xor(hi, off);
  return ((u128)hi << 64) | lo;
}

__device__ void block_reduce_store(unsigned long long cnt, u128 s, u128 s2, uint32_t mxb,
                                   uint32_t flags, StatsPartial* __restrict__ dst) {
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off);
    s += shfl_xor_u128(s, off);
    s2 += shfl_xor_u128(s2, off);
    mxb = max(mxb, (uint32_t)__shfl_xor((int)mxb, off));
    flags |= (uint32_t)__shfl_xor((int)flags, off);
  }
  __shared__ unsigned long long r_cnt[16];
  __shared__ u128 r_s[16], r_s2[16];
  __shared__ uint32_t r_mx[16], r_fl[16];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    r_cnt[wv] = cnt;
    r_s[wv] = s;
    r_s2[wv] = s2;
    r_mx[wv] = mxb;
    r_fl[wv] = flags;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nwv = blockDim.x >> 6;
    for (uint32_t i = 1; i < nwv; ++i) {
      cnt += r_cnt[i];
      s += r_s[i];
      s2 += r_s2[i];
      mxb = max(mxb, r_mx[i]);
      flags |= r_fl[i];
    }
    dst->count = cnt;
    dst->sum_lo = (unsigned long long)s;
    dst->sum_hi = (unsigned long long)(s >> 64);
    dst->sq_lo = (unsigned long long)s2;
    dst->sq_hi = (unsigned long long)(s2 >> 64);
    dst->max_bits = mxb;
    dst->flags = flags;
  }
}

// Clean-up passes 1-3 (src/local_thickness.rs:210-281), phase mask (:31-43)
// and the value part of nonzero_stats (src/metrics.rs:215-226), fused.  Only
// selected voxels (d2 > 0) carry a value.  A surface voxel's value is the f32
// mean, in NEIGHBORS_26 order, of its positive non-surface neighbours, or its
// own diameter when there is none.  Pass 2 only reads pass-1 results (surface
// values become negative and are never summed), so it is order-independent.
__global__ void __launch_bounds__(256)
k_cleanup_stats(const uint32_t* __restrict__ M, const uint32_t* __restrict__ d2,
                const float* __restrict__ diam, const PhaseStatus* __restrict__ st, uint32_t W,
                uint32_t H, uint32_t D, float scale_f32, int lsb_exp, float* __restrict__ map_out,
                StatsPartial* __restrict__ partials) {
  const uint64_t n = (uint64_t)W * H * D;
  const bool valid = st->has_site != 0;
  unsigned long long cnt = 0;
  u128 s = 0, s2 = 0;
  uint32_t mxb = 0, flags = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
    float t = 0.0f;
    if (valid && d2[idx] != 0) {
      const int x = (int)(idx % W);
      const uint64_t q = idx / W;
      const int y = (int)(q % H), z = (int)(q / H);
      const uint32_t m = M[idx];
      if (!surface_voxel(M, x, y, z, (int)W, (int)H, (int)D)) {
        t = diam[m];
      } else {
        float sum = 0.0f;
        uint32_t count = 0;
        for (int k = 0; k < 26; ++k) {
          const int nx = x + c_nb26[k][0], ny = y + c_nb26[k][1], nz = z + c_nb26[k][2];
          if (nx < 0 || ny < 0 || nz < 0 || nx >= (int)W || ny >= (int)H || nz >= (int)D) continue;
          const uint32_t mn = M[((size_t)nz * H + ny) * W + nx];
          if (mn == 0 || surface_voxel(M, nx, ny, nz, (int)W, (int)H, (int)D)) continue;
          sum += diam[mn];
          count += 1;
        }
        t = count > 0 ? sum / (float)count : diam[m];
      }
      if (t > 0.0f && t < INFINITY) acc_value(t * scale_f32, lsb_exp, cnt, s, s2, mxb, flags);
    }
    if (map_out) map_out[idx] = t;
  }
  block_reduce_store(cnt, s, s2, mxb, flags, partials + blockIdx.x);
}

__global__ void __launch_bounds__(256)
k_stats_final(const StatsPartial* __restrict__ partials, uint32_t count,
              StatsPartial* __restrict__ out) {
  unsigned long long cnt = 0;
  u128 s = 0, s2 = 0;
  uint32_t mxb = 0, flags = 0;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
    const StatsPartial p = partials[i];
    cnt += p.count;
    s += ((u128)p.sum_hi << 64) | p.sum_lo;
    s2 += ((u128)p.sq_hi << 64) | p.sq_lo;
    mxb = max(mxb, p.max_bits);
    flags |= p.flags;
  }
  block_reduce_store(cnt, s, s2, mxb, flags, out);
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t* __restrict__ p, uint64_t n, uint8_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void __launch_bounds__(256)
k_ridge_flags(const uint2* __restrict__ list, const PhaseStatus* __restrict__ st,
              uint8_t* __restrict__ flags) {
  const uint64_t count = st->ridge_count;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    flags[list[i].x] = 1;
}

// phase_inside (src/volume.rs:97-102): u8(roi != 0 && (bone != 0) == select_bone).
__global__ void __launch_bounds__(256)
k_phase_inside(const uint8_t* __restrict__ bone, const uint8_t* __restrict__ roi, uint64_t n,
               uint32_t select_bone, uint8_t* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = selected_voxel(bone[i], roi[i], select_bone) ? 1 : 0;
}

}  // namespace bm
