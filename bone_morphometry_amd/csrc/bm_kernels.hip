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
//
// Every stage kernel serves a list of items (ItemDesc): a block finds its item
// and its block within the item from the stage's start table; kernels with one
// block per item take the item from blockIdx.x.
#include <hip/hip_runtime.h>

#include "bm_device.h"
#include "bm_engine.h"

namespace bm {

// ---------------------------------------------------------------------------
// X pass.  Sites bits of 8 consecutive voxels x..x+7 of a row; bit j set when
// voxel x+j exists and is not selected.  Counts selected voxels.
// 0x80 in each zero byte of x, 0 elsewhere (no carry crosses a byte: the low
// seven bits of a byte plus 0x7F stay below 0x100).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}
// Bits 0..3 from bits 7, 15, 23 and 31 of m (m holds no other bits).
__device__ __forceinline__ uint32_t byte_msbs(uint32_t m) {
  return ((m >> 7) | (m >> 14) | (m >> 21) | (m >> 28)) & 0xFu;
}
__device__ __forceinline__ uint32_t site_bits8(BM_GLOBAL const uint8_t* __restrict__ data,
                                               BM_GLOBAL const uint8_t* __restrict__ roi, uint32_t want,
                                               size_t rb, uint32_t x, uint32_t W,
                                               uint32_t& selcnt) {
  uint32_t bits = 0;
  const bool aligned = ((reinterpret_cast<uintptr_t>(data + rb + x) |
                         (roi ? reinterpret_cast<uintptr_t>(roi + rb + x) : 0)) & 7) == 0;
  if (x + 8 <= W && aligned) {
    const v2u32 va = *(BM_GLOBAL const v2u32*)(data + rb + x);
    const uint2 v = make_uint2(va.x, va.y);
    uint2 r = make_uint2(0x01010101u, 0x01010101u);
    if (roi) {
      const v2u32 ra = *(BM_GLOBAL const v2u32*)(roi + rb + x);
      r = make_uint2(ra.x, ra.y);
    }
    // per byte (bit 7): selected = ROI byte nonzero and data byte nonzero as wanted
    const uint32_t zv0 = zero_bytes(v.x), zv1 = zero_bytes(v.y);
    const uint32_t sel0 = ~zero_bytes(r.x) & (want ? ~zv0 : zv0) & 0x80808080u;
    const uint32_t sel1 = ~zero_bytes(r.y) & (want ? ~zv1 : zv1) & 0x80808080u;
    selcnt += __popc(sel0) + __popc(sel1);
    bits = byte_msbs(sel0 ^ 0x80808080u) | (byte_msbs(sel1 ^ 0x80808080u) << 4);
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
constexpr uint32_t kXSegs = 4;  // rows of up to 2048 voxels take one sweep

// Zeroes the status (and the counters the later stages accumulate into) of
// each item; one block per item.  A kernel, not a memset: graph-safe.
__global__ void __launch_bounds__(256) k_prepare(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[blockIdx.x];
  uint32_t* st = reinterpret_cast<uint32_t*>(I.status);
  for (uint32_t i = threadIdx.x; i < sizeof(PhaseStatus) / 4; i += blockDim.x) st[i] = 0;
  if (I.flags & kItemThickness) {
    if (threadIdx.x == 0) {
      gbl(I.overflow)[0] = 0;
      gbl(I.bigtiles)[0] = 0;
    }
    if (I.flags & kItemQuantiles)
      for (uint32_t i = threadIdx.x; i < (uint32_t)kTopBins; i += blockDim.x) gbl(I.top16)[i] = 0;
  }
}

__global__ void __launch_bounds__(256)
k_edt_x(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts, uint32_t nitems) {
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const uint8_t* __restrict__ data = gbl(I.data);
  BM_GLOBAL const uint8_t* __restrict__ roi = gbl(I.roi);
  const uint32_t want = I.want_nonzero, W = I.W, rows = I.H * I.D;
  BM_GLOBAL uint16_t* __restrict__ out = gbl(I.dx);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gw = (vb * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (I.xblocks * blockDim.x) >> 6;
  const uint32_t nseg = (W + 511) / 512;
  uint32_t selcnt = 0, anysite = 0;
  __shared__ int s_carry[4][kXSegs];  // per wavefront: nearest site right of each segment
  // One segment (512 voxels, 8 per lane) of a row: each voxel's distance to
  // the nearest site on its left (in lower lanes of the segment, else carryL)
  // and on its right (in higher lanes, else carryR), as u16.
  auto segment_out = [&](size_t rb, uint32_t x0, uint32_t bits, int& carryL, int carryR) {
    const unsigned long long has = __ballot(bits != 0);
    const int last = bits ? (int)(x0 + 31 - __clz(bits)) : kNoneL;
    const int first = bits ? (int)(x0 + __ffs(bits) - 1) : kNoneR;
    const unsigned long long lower = has & ((1ull << lane) - 1ull);
    const int fromL = __shfl(last, lower ? 63 - __clzll(lower) : 0);
    const int Lc = lower ? fromL : carryL;
    const unsigned long long higher = lane == 63 ? 0ull : (has & (~0ull << (lane + 1)));
    const int fromR = __shfl(first, higher ? __ffsll(higher) - 1 : 0);
    const int Rc = higher ? fromR : carryR;
    if (has) carryL = __shfl(last, 63 - __clzll(has));
    if (x0 < W) {
      // nearest site at or left of x0 + j, and at or right of it, by one
      // step per voxel from Lc and from Rc
      int Ls[8];
      int Lrun = Lc, Rrun = Rc;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        Lrun = (bits >> j) & 1u ? (int)x0 + j : Lrun;
        Ls[j] = Lrun;
      }
      uint16_t o[8];
#pragma unroll
      for (int j = 7; j >= 0; --j) {
        Rrun = (bits >> j) & 1u ? (int)x0 + j : Rrun;
        const int x = (int)x0 + j;
        const int dist = min(x - Ls[j], Rrun - x);
        o[j] = dist > 65534 ? (uint16_t)kInf16 : (uint16_t)dist;
      }
      auto p = out + rb + x0;
      if (x0 + 8 <= W && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
        v4u32 w;
        w.x = o[0] | ((uint32_t)o[1] << 16);
        w.y = o[2] | ((uint32_t)o[3] << 16);
        w.z = o[4] | ((uint32_t)o[5] << 16);
        w.w = o[6] | ((uint32_t)o[7] << 16);
        *(BM_GLOBAL v4u32*)(p) = w;
      } else {
        for (int j = 0; j < 8; ++j)
          if (x0 + j < W) p[j] = o[j];
      }
    }
  };
  for (uint32_t r = gw; r < rows; r += nw) {
    const size_t rb = (size_t)r * W;
    int carryL = kNoneL;
    if (nseg > 1 && nseg <= kXSegs) {
      // Rows of up to kXSegs segments: the nearest site right of each segment
      // comes from a sweep from the right end that only reads (the first
      // site of each segment, a wave-uniform value), then each voxel is
      // written once; the second read of the row's bytes hits the cache.
      int* carryR = s_carry[threadIdx.x >> 6];
      int cr = kNoneR;
#pragma unroll 1
      for (uint32_t s = nseg; s-- > 0;) {
        if (lane == 0) carryR[s] = cr;
        uint32_t dummy = 0;
        const uint32_t x0 = s * 512 + lane * 8;
        const uint32_t bits = site_bits8(data, roi, want, rb, x0, W, dummy);
        const unsigned long long has = __ballot(bits != 0);
        const int first = bits ? (int)(x0 + __ffs(bits) - 1) : kNoneR;
        if (has) cr = __builtin_amdgcn_readlane(first, __ffsll(has) - 1);
      }
      // (carryR is written and read by this wavefront only)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
      for (uint32_t s = 0; s < nseg; ++s) {
        const uint32_t x0 = s * 512 + lane * 8;
        const uint32_t bits = site_bits8(data, roi, want, rb, x0, W, selcnt);
        anysite |= bits;
        segment_out(rb, x0, bits, carryL, carryR[s]);
      }
      continue;
    }
    for (uint32_t s = 0; s < nseg; ++s) {
      const uint32_t x0 = s * 512 + lane * 8;
      const uint32_t bits = site_bits8(data, roi, want, rb, x0, W, selcnt);
      anysite |= bits;
      segment_out(rb, x0, bits, carryL, kNoneR);
    }
    if (nseg == 1) continue;
    // Longer rows: sites right of each segment from a second sweep from the
    // right end.
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
        auto p = out + rb + x0;
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
  // Per-block partials (reduced by k_reduce_status): no contended atomics.
  const unsigned long long tot = wave_sum_u64(selcnt);
  const unsigned long long anyb = __ballot(anysite != 0);
  __shared__ unsigned long long s_sel[4];
  __shared__ uint32_t s_any[4];
  if (lane == 0) {
    s_sel[threadIdx.x >> 6] = tot;
    s_any[threadIdx.x >> 6] = anyb != 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0;
    uint32_t b = 0;
    for (uint32_t i = 0; i < (blockDim.x >> 6); ++i) {
      a += s_sel[i];
      b |= s_any[i];
    }
    gbl(I.xpart)[vb].selected = a;
    gbl(I.xpart)[vb].flag = b;
  }
}

// ---------------------------------------------------------------------------
// The 1-D step of the Y and Z passes for one lane: d(q) = min_p (q-p)^2 + f(p)
// for the ROWS consecutive rows q0 .. q0 + ROWS - 1 below end.  At offset o the candidates of row k are q0+k-o and q0+k+o, a window
// that slides by one row per offset.  Offsets go in blocks of ROWS: a block
// reads the ROWS new values on each side once, and every candidate of the
// block is then a compile-time element of the previous or the new values, so
// the window slides without register moves.  The search stops once o^2
// reaches every row's best, which is exact: every later candidate adds at
// least o^2.  The stop is taken when it holds for every lane of the
// wavefront, so the offset and all window positions are the same across the
// wavefront; a lane that could have stopped earlier only adds candidates that
// cannot lower its minima.  f(p) is kInf32 outside [0, n) (kInf32 + o^2 stays
// below 2^31 for every offset reached); n bounds the offsets when no finite
// candidate exists.  fin(p) reads the staged tile without checks and is used
// when a whole block's positions lie in [in_lo, in_hi).
constexpr int kEdtRows = 8;
constexpr int kStageLoads = 4;  // Y/Z tile staging: items per thread with loads in flight
template <int ROWS, class F, class Fin, class Out>
__device__ __forceinline__ void window_group(uint32_t q0, uint32_t end, uint32_t n, int in_lo,
                                             int in_hi, F&& f, Fin&& fin, Out&& out) {
  const int iq = (int)q0;
  uint32_t best[ROWS], lp[ROWS], up[ROWS];
  if (__ballot(!(iq >= in_lo && iq + ROWS <= in_hi)) == 0) {
#pragma unroll
    for (int k = 0; k < ROWS; ++k) best[k] = fin(iq + k);
  } else {
#pragma unroll
    for (int k = 0; k < ROWS; ++k) best[k] = f(iq + k);
  }
  uint32_t bmax = 0;
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    lp[k] = up[k] = best[k];  // f at q0 + k: the low side at o0 = 0, and the high side
    if (q0 + k < end) bmax = max(bmax, best[k]);
  }
  for (uint32_t o0 = 0;; o0 += ROWS) {  // offsets o0 + 1 .. o0 + ROWS
    if (!__ballot((o0 + 1) * (o0 + 1) < bmax)) break;
    // nothing left on either side of any row of any lane
    if (!__ballot(!(iq + ROWS - 1 < (int)(o0 + 1) && q0 + o0 + 1 >= n))) break;
    // ln[m] = f(q0 - o0 - ROWS + m), un[m] = f(q0 + o0 + ROWS + m); lp and up
    // hold f(q0 - o0 + m) and f(q0 + o0 + m)
    const int plo = iq - (int)o0 - ROWS, phi = iq + (int)o0 + ROWS;
    uint32_t ln[ROWS], un[ROWS];
    if (__ballot(!(plo >= in_lo && phi + ROWS <= in_hi)) == 0) {
#pragma unroll
      for (int m = 0; m < ROWS; ++m) {
        ln[m] = fin(plo + m);
        un[m] = fin(phi + m);
      }
    } else {
#pragma unroll
      for (int m = 0; m < ROWS; ++m) {
        ln[m] = f(plo + m);
        un[m] = f(phi + m);
      }
    }
    bool stop = false;
#pragma unroll
    for (int j = 1; j <= ROWS; ++j) {
      const uint32_t o = o0 + j, oo = o * o;
      bmax = 0;
#pragma unroll
      for (int k = 0; k < ROWS; ++k) {
        const int ci = ROWS + k - j;  // q0 + k - o = plo + ci
        const int di = k + j;         // q0 + k + o = q0 + o0 + di
        const uint32_t lo = ci < ROWS ? ln[ci] : lp[ci - ROWS];
        const uint32_t hi = di < ROWS ? up[di] : un[di - ROWS];
        best[k] = min(best[k], min(lo, hi) + oo);
        if (q0 + k < end) bmax = max(bmax, best[k]);
      }
      if (j < ROWS && !__ballot((o + 1) * (o + 1) < bmax)) {
        stop = true;
        break;
      }
    }
    if (stop) break;
#pragma unroll
    for (int m = 0; m < ROWS; ++m) {
      lp[m] = ln[m];
      up[m] = un[m];
    }
  }
#pragma unroll
  for (int k = 0; k < ROWS; ++k)
    if (q0 + k < end) out(q0 + k, best[k]);
}

// The same search on saturated 16-bit values, two rows per 32-bit register:
// pair (p, p + 1), p even, holds min(f(p), 0xFFFF) in the low half and
// min(f(p + 1), 0xFFFF) in the high half, and one packed minimum or
// saturating add serves two rows.  Saturation is monotone, so a row's packed
// result is min(d(q), 0xFFFF) and is exact below 0xFFFF; a saturated row
// makes the call return false without writing, and the caller runs the
// 32-bit search for the group.  Offsets stop by o = 255 at the latest, where
// o^2 alone saturates.  Rows at or past end start at 0, so they never hold
// the search open and are never written.  The candidates of two rows at an
// odd offset straddle two pairs: those pairs come from one byte alignment
// per pair and block of offsets.  q0 must be even; f2 and fin2 take an even
// position and return its pair (fin2 from the staged tile when a whole
// block lies in [in_lo, in_hi)).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_min16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_add_sat16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2, a),
                                                                    __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t sat16(uint32_t v) { return min(v, 0xFFFFu); }
__device__ __forceinline__ uint32_t pair16(uint32_t lo, uint32_t hi) {
  return sat16(lo) | (sat16(hi) << 16);
}

template <int ROWS, class F2, class Fin2, class Out>
__device__ __forceinline__ bool window_group16(uint32_t q0, uint32_t end, uint32_t n, int in_lo,
                                               int in_hi, F2&& f2, Fin2&& fin2, Out&& out) {
  constexpr int P = ROWS / 2;
  static_assert(ROWS % 2 == 0, "rows go in pairs");
  const int iq = (int)q0;
  uint32_t best[P], lp[P], up[P];
  if (__ballot(!(iq >= in_lo && iq + ROWS <= in_hi)) == 0) {
#pragma unroll
    for (int i = 0; i < P; ++i) best[i] = fin2(iq + 2 * i);
  } else {
#pragma unroll
    for (int i = 0; i < P; ++i) best[i] = f2(iq + 2 * i);
  }
  uint32_t bm = 0;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    lp[i] = up[i] = best[i];  // pairs at q0 + 2i: the low side at o0 = 0, and the high side
    const uint32_t r = q0 + 2 * i;
    best[i] &= (r < end ? 0xFFFFu : 0u) | (r + 1 < end ? 0xFFFF0000u : 0u);
    bm = pk_max16(bm, best[i]);
  }
  uint32_t bmax = max(bm & 0xFFFFu, bm >> 16);
  for (uint32_t o0 = 0;; o0 += ROWS) {  // offsets o0 + 1 .. o0 + ROWS
    if (!__ballot((o0 + 1) * (o0 + 1) < bmax)) break;
    // nothing left on either side of any row of any lane
    if (!__ballot(!(iq + ROWS - 1 < (int)(o0 + 1) && q0 + o0 + 1 >= n))) break;
    // lw = pairs at q0 - o0 - ROWS + 2m (ln) then q0 - o0 + 2m (lp); uw = pairs
    // at q0 + o0 + 2m (up) then q0 + o0 + ROWS + 2m (un)
    const int plo = iq - (int)o0 - ROWS, phi = iq + (int)o0 + ROWS;
    uint32_t lw[2 * P], uw[2 * P];
    if (__ballot(!(plo >= in_lo && phi + ROWS <= in_hi)) == 0) {
#pragma unroll
      for (int m = 0; m < P; ++m) {
        lw[m] = fin2(plo + 2 * m);
        uw[P + m] = fin2(phi + 2 * m);
      }
    } else {
#pragma unroll
      for (int m = 0; m < P; ++m) {
        lw[m] = f2(plo + 2 * m);
        uw[P + m] = f2(phi + 2 * m);
      }
    }
#pragma unroll
    for (int m = 0; m < P; ++m) {
      lw[P + m] = lp[m];
      uw[m] = up[m];
    }
    // the pairs that straddle: (high half of pair m, low half of pair m + 1)
    uint32_t ls[2 * P - 1], us[2 * P - 1];
#pragma unroll
    for (int m = 0; m < 2 * P - 1; ++m) {
      ls[m] = __builtin_amdgcn_alignbyte(lw[m + 1], lw[m], 2);
      us[m] = __builtin_amdgcn_alignbyte(uw[m + 1], uw[m], 2);
    }
    bool stop = false;
#pragma unroll
    for (int j = 1; j <= ROWS; ++j) {
      const uint32_t o = o0 + j, oo2 = sat16(o * o) * 0x10001u;
      bm = 0;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        // rows 2i and 2i + 1: low candidates at ROWS + 2i - j of lw's rows,
        // high candidates at 2i + j of uw's rows
        const int c = ROWS + 2 * i - j, d = 2 * i + j;
        const uint32_t lo = (j & 1) ? ls[(c - 1) / 2] : lw[c / 2];
        const uint32_t hi = (j & 1) ? us[(d - 1) / 2] : uw[d / 2];
        best[i] = pk_min16(best[i], pk_add_sat16(pk_min16(lo, hi), oo2));
        bm = pk_max16(bm, best[i]);
      }
      bmax = max(bm & 0xFFFFu, bm >> 16);
      if (j < ROWS && !__ballot((o + 1) * (o + 1) < bmax)) {
        stop = true;
        break;
      }
    }
    if (stop) break;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      lp[m] = lw[m];
      up[m] = uw[P + m];
    }
  }
  bool sat = false;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const uint32_t r = q0 + 2 * i;
    sat |= (r < end && (best[i] & 0xFFFFu) == 0xFFFFu) || (r + 1 < end && (best[i] >> 16) == 0xFFFFu);
  }
  if (__ballot(sat)) return false;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const uint32_t r = q0 + 2 * i;
    if (r < end) out(r, best[i] & 0xFFFFu);
    if (r + 1 < end) out(r + 1, best[i] >> 16);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Y pass: tile of 64 x-columns x the rows [ya, ya + 2 np) of one z-slice in
// LDS as saturated 16-bit squares, one word per column and pair of rows (the
// tile starts on an even row: chunks are even, so y0 is).
__global__ void __launch_bounds__(256)
k_edt_y(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts, uint32_t nitems) {
  extern __shared__ uint32_t ytile[];
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const uint16_t* __restrict__ in = gbl(I.dx);
  BM_GLOBAL uint32_t* __restrict__ out = gbl(I.h);
  const uint32_t W = I.W, H = I.H, chunk = I.ychunk, halo = I.yhalo;
  const uint32_t bx = vb % I.ygx, bz = vb / I.ygx / I.ygy;
  const uint32_t x0 = bx * 64, z = (vb / I.ygx) % I.ygy;
  const uint32_t y0 = bz * chunk, y1 = min(H, y0 + chunk);
  const uint32_t ya = (y0 > halo ? y0 - halo : 0) & ~1u, yb = min(H, y1 + halo);
  const uint32_t np = (yb - ya + 1) / 2;
  const size_t sbase = (size_t)z * W * H;
  const bool vec = (x0 + 64 <= W) && ((W & 7) == 0);
  auto put = [&](uint32_t pr, uint32_t c, const uint32_t* a, const uint32_t* b) {
    uint32_t w[8];
    for (int k = 0; k < 8; ++k) w[k] = pair16(sq16(a[k]), sq16(b[k]));
    *reinterpret_cast<uint4*>(&ytile[pr * 64 + c]) = make_uint4(w[0], w[1], w[2], w[3]);
    *reinterpret_cast<uint4*>(&ytile[pr * 64 + c + 4]) = make_uint4(w[4], w[5], w[6], w[7]);
  };
  const uint32_t nitem = np * 8;  // (pair of rows, 8 columns)
  if (vec) {
    // kStageLoads items' loads in flight per thread before the first is used
    for (uint32_t i0 = threadIdx.x; i0 < nitem; i0 += kStageLoads * blockDim.x) {
      v4u32 w0[kStageLoads], w1[kStageLoads];
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const uint32_t i = i0 + u * blockDim.x, r0 = ya + 2 * (i >> 3);  // r0 < H
        auto p = in + sbase + (size_t)r0 * W + x0 + (i & 7) * 8;
        w0[u] = w1[u] = v4u32{~0u, ~0u, ~0u, ~0u};
        if (i < nitem) {
          w0[u] = *(BM_GLOBAL const v4u32*)(p);
          if (r0 + 1 < H) w1[u] = *(BM_GLOBAL const v4u32*)(p + W);
        }
      }
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const uint32_t i = i0 + u * blockDim.x;
        if (i >= nitem) break;
        uint32_t a[8], b[8];
        for (int k = 0; k < 4; ++k) {
          a[2 * k] = w0[u][k] & 0xFFFFu;
          a[2 * k + 1] = w0[u][k] >> 16;
          b[2 * k] = w1[u][k] & 0xFFFFu;
          b[2 * k + 1] = w1[u][k] >> 16;
        }
        put(i >> 3, (i & 7) * 8, a, b);
      }
    }
  } else {
    for (uint32_t i = threadIdx.x; i < nitem; i += blockDim.x) {
      const uint32_t pr = i >> 3, c = (i & 7) * 8, r0 = ya + 2 * pr;  // r0 < H
      auto p = in + sbase + (size_t)r0 * W + x0 + c;
      const bool has1 = r0 + 1 < H;
      uint32_t a[8], b[8];
      for (int k = 0; k < 8; ++k) {
        const bool in_row = x0 + c + k < W;
        a[k] = in_row ? p[k] : kInf16;
        b[k] = in_row && has1 ? p[W + k] : kInf16;
      }
      put(pr, c, a, b);
    }
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const uint32_t x = x0 + lane;
  if (x >= W) return;
  auto col = in + sbase + x;
  // f(p) = dx(p)^2 of this column, kInf32 outside it, from global memory
  auto fy = [&](int p) -> uint32_t {
    if (p < 0 || p >= (int)H) return kInf32;
    return sq16(col[(size_t)p * W]);
  };
  // the pair at even p: from the tile, else from global memory
  const int tlo = (int)ya, thi = (int)(ya + 2 * np);
  auto fin2 = [&](int p) -> uint32_t { return ytile[(uint32_t)(p - tlo) / 2 * 64 + lane]; };
  auto f2 = [&](int p) -> uint32_t {
    const uint32_t r = (uint32_t)(p - tlo);
    if (r < 2 * np) return ytile[r / 2 * 64 + lane];
    return pair16(fy(p), fy(p + 1));
  };
  auto outy = [&](uint32_t q, uint32_t v) { out[sbase + (size_t)q * W + x] = v; };
  for (uint32_t q0 = y0 + kEdtRows * wv; q0 < y1; q0 += kEdtRows * nwv)
    if (!window_group16<kEdtRows>(q0, y1, H, tlo, thi, f2, fin2, outy))
      window_group<kEdtRows>(q0, y1, H, 0, 0, fy, fy, outy);
}

// ---------------------------------------------------------------------------
// Z pass: tile of 32 x-columns x the planes [za, za + 2 np) at one y in LDS as
// saturated 16-bit values, one word per column and pair of planes (za even).
__global__ void __launch_bounds__(256)
k_edt_z(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts, uint32_t nitems) {
  extern __shared__ uint32_t ztile[];
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const uint32_t* __restrict__ in = gbl(I.h);
  BM_GLOBAL uint32_t* __restrict__ out = gbl(I.d2);
  const uint32_t W = I.W, H = I.H, D = I.D, chunk = I.zchunk, halo = I.zhalo;
  const uint32_t x0 = (vb % I.zgx) * 32, y = (vb / I.zgx) % I.zgy;
  const uint32_t z0 = (vb / I.zgx / I.zgy) * chunk, z1 = min(D, z0 + chunk);
  const uint32_t za = (z0 > halo ? z0 - halo : 0) & ~1u, zb = min(D, z1 + halo);
  const uint32_t np = (zb - za + 1) / 2;
  const size_t slice = (size_t)W * H;
  const size_t base = (size_t)y * W + x0;
  const bool vec = (x0 + 32 <= W) && ((W & 3) == 0);
  const uint32_t nitem = np * 8;  // (pair of planes, 4 columns)
  auto put = [&](uint32_t i, const uint32_t* a, const uint32_t* b) {
    *reinterpret_cast<uint4*>(&ztile[(i >> 3) * 32 + (i & 7) * 4]) =
        make_uint4(pair16(a[0], b[0]), pair16(a[1], b[1]), pair16(a[2], b[2]), pair16(a[3], b[3]));
  };
  if (vec) {
    // kStageLoads items' loads in flight per thread before the first is used
    for (uint32_t i0 = threadIdx.x; i0 < nitem; i0 += kStageLoads * blockDim.x) {
      v4u32 w0[kStageLoads], w1[kStageLoads];
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const uint32_t i = i0 + u * blockDim.x, r0 = za + 2 * (i >> 3);  // r0 < D
        auto p = in + (size_t)r0 * slice + base + (i & 7) * 4;
        w0[u] = w1[u] = v4u32{kInf32, kInf32, kInf32, kInf32};
        if (i < nitem) {
          w0[u] = *(BM_GLOBAL const v4u32*)(p);
          if (r0 + 1 < D) w1[u] = *(BM_GLOBAL const v4u32*)(p + slice);
        }
      }
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const uint32_t i = i0 + u * blockDim.x;
        if (i >= nitem) break;
        const uint32_t a[4] = {w0[u].x, w0[u].y, w0[u].z, w0[u].w};
        const uint32_t b[4] = {w1[u].x, w1[u].y, w1[u].z, w1[u].w};
        put(i, a, b);
      }
    }
  } else {
    for (uint32_t i = threadIdx.x; i < nitem; i += blockDim.x) {
      const uint32_t c = (i & 7) * 4, r0 = za + 2 * (i >> 3);  // r0 < D
      auto p = in + (size_t)r0 * slice + base + c;
      const bool has1 = r0 + 1 < D;
      uint32_t a[4], b[4];
      for (int k = 0; k < 4; ++k) {
        const bool in_row = x0 + c + k < W;
        a[k] = in_row ? p[k] : kInf32;
        b[k] = in_row && has1 ? p[slice + k] : kInf32;
      }
      put(i, a, b);
    }
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const uint32_t xl = lane & 31, zs = lane >> 5;
  const uint32_t x = x0 + xl;
  uint32_t mx = 0;
  if (x < W) {
    auto col = in + base + xl;
    auto fz = [&](int p) -> uint32_t {  // kInf32 outside the line, from global memory
      if (p < 0 || p >= (int)D) return kInf32;
      return col[(size_t)p * slice];
    };
    const int tlo = (int)za;
    auto fin2 = [&](int p) -> uint32_t { return ztile[(uint32_t)(p - tlo) / 2 * 32 + xl]; };
    auto f2 = [&](int p) -> uint32_t {
      const uint32_t r = (uint32_t)(p - tlo);
      if (r < 2 * np) return ztile[r / 2 * 32 + xl];
      return pair16(fz(p), fz(p + 1));
    };
    auto outz = [&](uint32_t q, uint32_t v) {
      out[(size_t)q * slice + base + xl] = v;
      mx = max(mx, v);
    };
    // the two half-waves take adjacent groups of rows
    for (uint32_t q0 = z0 + 2 * kEdtRows * wv + kEdtRows * zs; q0 < z1; q0 += 2 * kEdtRows * nwv)
      if (!window_group16<kEdtRows>(q0, z1, D, tlo, (int)(za + 2 * np), f2, fin2, outz))
        window_group<kEdtRows>(q0, z1, D, 0, 0, fz, fz, outz);
  }
  mx = wave_max_u32(mx);
  __shared__ uint32_t s_mx[4];
  if (lane == 0) s_mx[wv] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < nwv; ++i) r = max(r, s_mx[i]);
    gbl(I.zmax)[vb] = r;
  }
}

// Status of a phase from the EDT passes' per-block partials (one block).
__global__ void __launch_bounds__(1024) k_reduce_status(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[blockIdx.x];
  BM_GLOBAL const BlockPartial* __restrict__ xpart = gbl(I.xpart);
  BM_GLOBAL const uint32_t* __restrict__ zmax = gbl(I.zmax);
  const uint32_t nx = I.xblocks, nz = I.zgx * I.zgy * I.zgz;
  BM_GLOBAL PhaseStatus* __restrict__ st = gbl(I.status);
  unsigned long long sel = 0;
  uint32_t any = 0, mx = 0;
  for (uint32_t i = threadIdx.x; i < nx; i += blockDim.x) {
    sel += xpart[i].selected;
    any |= xpart[i].flag;
  }
  for (uint32_t i = threadIdx.x; i < nz; i += blockDim.x) mx = max(mx, zmax[i]);
  sel = wave_sum_u64(sel);
  mx = wave_max_u32(mx);
  any = __ballot(any != 0) != 0;
  __shared__ unsigned long long s_sel[16];
  __shared__ uint32_t s_mx[16], s_any[16];
  if ((threadIdx.x & 63) == 0) {
    s_sel[threadIdx.x >> 6] = sel;
    s_mx[threadIdx.x >> 6] = mx;
    s_any[threadIdx.x >> 6] = any;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t i = 1; i < (blockDim.x >> 6); ++i) {
      sel += s_sel[i];
      mx = max(mx, s_mx[i]);
      any |= s_any[i];
    }
    st->selected = sel;
    st->has_site = any;
    st->max_d2 = mx;
  }
}

// Templates T_c[r^2] (scan_cube, src/local_thickness.rs:75-102, for the
// offsets (1,0,0), (1,1,0), (1,1,1)) and diameters for every r^2 in
// [0, max_d2].  One wavefront per r^2: lanes take k = lane, lane + 64, ...;
// along j the integer square root isqrt(r^2 - k^2 - j^2) only decreases, so it
// is stepped down instead of recomputed.
__global__ void __launch_bounds__(256)
k_luts(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts, uint32_t nitems) {
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const PhaseStatus* __restrict__ st = gbl(I.status);
  if (!st->has_site) return;
  const uint32_t mx = st->max_d2;
  const uint32_t lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
  const uint32_t nw = I.lblocks * wpb, L = I.bound + 1;
  for (uint32_t m = vb * wpb + (threadIdx.x >> 6); m <= mx; m += nw) {
    if (lane == 0) gbl(I.diam)[m] = diameter_of(m);
    if (m == 0) continue;
    const uint32_t radius = isqrt_u32(m) + 1;
    uint32_t t1 = 0, t2 = 0, t3 = 0;
    for (uint32_t k = lane; k <= radius; k += 64) {
      const uint32_t sk = k * k;
      if (sk > m) break;
      uint32_t i = isqrt_u32(m - sk);
      for (uint32_t j = 0;; ++j) {
        const uint32_t skj = sk + j * j;
        if (skj > m) break;
        while (i * i > m - skj) --i;
        const uint32_t ip = (i + 1) * (i + 1), j1 = (j + 1) * (j + 1);
        t1 = max(t1, skj + ip);                       // (k, j, i+1)
        t2 = max(t2, sk + j1 + ip);                   // (k, j+1, i+1)
        t3 = max(t3, (k + 1) * (k + 1) + j1 + ip);    // (k+1, j+1, i+1)
      }
    }
    t1 = wave_max_u32(t1);
    t2 = wave_max_u32(t2);
    t3 = wave_max_u32(t3);
    if (lane == 0) {
      gbl(I.tmpl)[m] = t1;
      gbl(I.tmpl)[L + m] = t2;
      gbl(I.tmpl)[2 * L + m] = t3;
    }
  }
}

// Exact statistics of an item from its clean-up partials; one block per item.
__global__ void __launch_bounds__(256) k_stats_final(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[blockIdx.x];
  if (!(I.flags & kItemThickness)) return;
  const StatsPartial* __restrict__ partials = I.partials;
  const uint32_t count = I.cblocks;
  StatsPartial* __restrict__ out = I.stats;
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

// Per-item results (status, statistics, selection) gathered into one array
// for a single copy to the host; one block per item.
__global__ void __launch_bounds__(64) k_collect(const ItemDesc* __restrict__ items,
                                               ItemResult* __restrict__ out) {
  const ItemDesc& I = items[blockIdx.x];
  uint32_t* r = reinterpret_cast<uint32_t*>(out + blockIdx.x);
  const uint32_t* st = reinterpret_cast<const uint32_t*>(I.status);
  const uint32_t* sp = (I.flags & kItemThickness) ? reinterpret_cast<const uint32_t*>(I.stats) : nullptr;
  const uint32_t* se = (I.flags & kItemQuantiles) ? reinterpret_cast<const uint32_t*>(I.sel) : nullptr;
  constexpr uint32_t n0 = sizeof(PhaseStatus) / 4, n1 = sizeof(StatsPartial) / 4,
                     n2 = sizeof(SelectState) / 4;
  static_assert(sizeof(ItemResult) == 4 * (n0 + n1 + n2), "ItemResult has no padding");
  for (uint32_t i = threadIdx.x; i < n0 + n1 + n2; i += blockDim.x) {
    uint32_t v = 0;
    if (i < n0) v = st[i];
    else if (i < n0 + n1) v = sp ? sp[i - n0] : 0u;
    else v = se ? se[i - n0 - n1] : 0u;
    r[i] = v;
  }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t* __restrict__ p, uint64_t n, uint8_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
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
