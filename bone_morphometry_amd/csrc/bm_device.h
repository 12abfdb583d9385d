// bm_device.h -- device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "bm_engine.h"

namespace bm {

typedef unsigned __int128 u128;

// Device pointers taken from an ItemDesc in memory are generic to the
// compiler, which then emits flat loads and stores; flat accesses also count
// against the LDS wait counter, so every LDS wait would wait for them too.
// gbl() states that a pointer addresses global memory.
#define BM_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ BM_GLOBAL T* gbl(T* p) {
  return (BM_GLOBAL T*)p;
}
// Vector types usable through global-address-space pointers.
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

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

// Diameter (2 * sqrt_f64(m)) as f32 (src/local_thickness.rs:206).  For
// m < 2^24, (float)m is exact and the correctly rounded f32 sqrt scaled by 2
// equals RN32(2 sqrt(m)); the reference's RN32(RN64(2 sqrt(m))) is the same
// value because 2 sqrt(m) never lies within 2^-53 (relative) of an f32
// rounding midpoint.  Larger m take the f64 route.
__device__ __forceinline__ float diameter_of(uint32_t m) {
  if (m < (1u << 24)) return 2.0f * sqrtf((float)m);
  return (float)(2.0 * sqrt((double)m));
}

// ---------------------------------------------------------------------------
// NEIGHBORS_26 in Fiji's visiting order (src/local_thickness.rs:283-312); the
// f32 sum of the clean-up runs in this order.
constexpr int8_t kNb26[26][3] = {
    {0, 0, -1},  {0, 0, 1},   {0, -1, 0},  {0, 1, 0},   {-1, 0, 0},  {1, 0, 0},  {0, 1, -1},
    {0, 1, 1},   {1, -1, 0},  {1, 1, 0},   {-1, 0, 1},  {1, 0, 1},   {0, -1, -1}, {0, -1, 1},
    {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1}, {1, 0, -1},  {1, 1, 1},   {1, -1, 1},  {-1, 1, 1},
    {-1, -1, 1}, {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};

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
  // (the shifts are 64-bit ones: mant << shift < 2^48, and mant^2 < 2^48 is
  // split at bit 64 - 2 shift; a 128-bit variable shift costs far more)
  cnt += 1;
  s += (u128)((unsigned long long)mant << shift);
  const unsigned long long q = (unsigned long long)mant * mant;
  const int n = 2 * shift;
  const unsigned long long lo = q << n, hi = n ? q >> (64 - n) : 0ull;
  s2 += ((u128)hi << 64) | lo;
  mxb = max(mxb, bits);
}

__device__ __forceinline__ u128 shfl_xor_u128(u128 v, int off) {
  unsigned long long lo = (unsigned long long)v, hi = (unsigned long long)(v >> 64);
  lo = __shfl_xor(lo, off);
  hi = __shfl_xor(hi, off);
  return ((u128)hi << 64) | lo;
}

__device__ inline void block_reduce_store(unsigned long long cnt, u128 s, u128 s2, uint32_t mxb,
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

// Inclusive sum over lanes 0..lane of a converged wavefront, without LDS:
// Hillis-Steele steps of 1, 2, 4 and 8 lanes inside each row of 16 (DPP row
// shifts; a lane with no source lane adds 0), then each row adds the totals of
// the rows below it (lanes 15, 31 and 47 read as scalars).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 15);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 31);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 47);
  const uint32_t lane = threadIdx.x & 63;
  return x + (lane >= 16 ? r0 : 0u) + (lane >= 32 ? r1 : 0u) + (lane >= 48 ? r2 : 0u);
}

// Adds 1 to hist[bin] for every active lane, one LDS atomic per distinct bin of
// the wavefront (thickness maps repeat values, so lanes often share a bin).
__device__ __forceinline__ void wave_hist_add(uint32_t* hist, bool active, uint32_t bin) {
  unsigned long long act = __ballot(active);
  while (act) {
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
    const unsigned long long same = __ballot(active && bin == v);
    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&hist[v], (uint32_t)__popcll(same));
    if (bin == v) active = false;
    act &= ~same;
  }
}

// The same counts without a loop: one LDS atomic per run of consecutive active
// lanes holding the same bin, the run's first lane adding the run's length.
// Suited to lanes that hold neighbouring voxels of a map (runs are long); the
// wavefront must be converged.
__device__ __forceinline__ void wave_hist_add_runs(uint32_t* hist, bool active, uint32_t bin) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t key = active ? bin : 0xFFFFFFFFu;
  // the key of the lane below (DPP wave shift; lane 0 gets 0xFFFFFFFF, no key)
  const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x138, 0xF, 0xF, false);
  const bool head = active && prev != key;
  // a run ends before the next head or the next inactive lane
  const unsigned long long stop = __ballot(head) | ~__ballot(active);
  if (head) {
    const unsigned long long rest = lane == 63 ? 0ull : stop >> (lane + 1);  // lanes above
    atomicAdd(&hist[bin], rest ? (uint32_t)__ffsll(rest) : 64u - lane);
  }
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8), each XCD with its own L2.  This maps block b to a tile so that each
// XCD walks one contiguous range of tiles: neighbouring tiles, whose halos and
// ridge lists overlap, then share an L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t ntiles) {
  const uint32_t per = (ntiles + kXcds - 1) / kXcds;
  if (b / kXcds >= per) return ntiles;  // more blocks than positions: none
  return (b % kXcds) * per + b / kXcds;  // >= ntiles in the last XCD's tail: callers skip
}

// Tile order of the tile kernels: 4x4x4 groups of 8^3 tiles (64 consecutive
// positions each, x fastest inside), the groups in x-fastest order.  A short
// run of positions then covers tiles that are neighbours along all three
// axes, whose halos and ridge lists overlap in L2.  Positions of a partial
// group beyond the volume give kNoTile.
// With grouped == false the order is plain x-fastest (position = tile).
constexpr uint32_t kNoTile = 0xFFFFFFFFu;
__host__ __device__ inline uint32_t tile_positions(uint32_t TX, uint32_t TY, uint32_t TZ,
                                                   bool grouped) {
  return grouped ? ((TX + 3) / 4) * ((TY + 3) / 4) * ((TZ + 3) / 4) * 64 : TX * TY * TZ;
}
// Exact quotient n / d of 32-bit unsigned numbers for a divisor d >= 1 fixed
// in advance: one high multiply and two shifts (Granlund and Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1).
struct FastDiv {
  uint32_t d, m, s1, s2;
  __host__ __device__ explicit FastDiv(uint32_t divisor) : d(divisor) {
    uint32_t l = 0;  // ceil(log2 d)
    while (l < 32 && (1ull << l) < (unsigned long long)divisor) ++l;
    m = (uint32_t)(((((1ull << l) - divisor) << 32) / divisor) + 1);
    s1 = l > 0 ? 1u : 0u;
    s2 = l > 0 ? l - 1 : 0u;
  }
  __host__ __device__ FastDiv(uint32_t d_, uint32_t m_, uint32_t s1_, uint32_t s2_)
      : d(d_), m(m_), s1(s1_), s2(s2_) {}
  __host__ __device__ uint32_t div(uint32_t n) const {
    const uint32_t t = (uint32_t)(((unsigned long long)m * n) >> 32);
    return (t + ((n - t) >> s1)) >> s2;
  }
};

// Coordinates of the tile at position j of that order, with invariant-divisor
// quotients; at() is false when the position holds no tile (kNoTile).
struct TileWalk {
  uint32_t TX, TY, TZ;
  bool grouped;
  FastDiv fx, fy;  // SX, SY (grouped) or TX, TY (plain)
  __device__ TileWalk(uint32_t tx_, uint32_t ty_, uint32_t tz_, bool g)
      : TX(tx_), TY(ty_), TZ(tz_), grouped(g), fx(g ? (tx_ + 3) >> 2 : tx_),
        fy(g ? (ty_ + 3) >> 2 : ty_) {}
  // with the divisor constants computed in advance (fx: TX or SX, fy: TY or SY)
  __device__ TileWalk(uint32_t tx_, uint32_t ty_, uint32_t tz_, bool g, FastDiv fx_, FastDiv fy_)
      : TX(tx_), TY(ty_), TZ(tz_), grouped(g), fx(fx_), fy(fy_) {}
  __device__ bool at(uint32_t j, uint32_t& tx, uint32_t& ty, uint32_t& tz) const {
    if (!grouped) {
      const uint32_t q = fx.div(j), r = fy.div(q);
      tx = j - q * TX;
      ty = q - r * TY;
      tz = r;
      return tz < TZ;
    }
    const uint32_t s = j >> 6, l = j & 63;
    const uint32_t q = fx.div(s), r = fy.div(q);
    tx = (s - q * fx.d) * 4 + (l & 3);
    ty = (q - r * fy.d) * 4 + ((l >> 2) & 3);
    tz = r * 4 + (l >> 4);
    return tx < TX && ty < TY && tz < TZ;
  }
};

// Item of block b in a stage: the last item whose start is <= b (items with
// no block in the stage share their start with the next one).  starts has
// nitems + 1 entries; the result and the block of the item are uniform over
// the block, so the search runs on scalar loads.
__device__ __forceinline__ uint32_t item_of(const uint32_t* __restrict__ starts, uint32_t nitems,
                                            uint32_t b, uint32_t& vb) {
  uint32_t lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  vb = b - starts[lo];
  return lo;
}

}  // namespace bm
