// bm_quantiles.hip -- radix select of the order statistics behind the five
// quantiles of nonzero_stats (src/metrics.rs:242-255,271-298), per item.
//
// The values are positive finite f32, whose bit patterns sort like the values,
// so the k-th smallest value is found digit by digit, most significant first.
// The clean-up kernel already counts the top 16 bits (kTopBase + bin); two
// 8-bit passes over the map follow.  Each pass histograms, for every wanted
// rank, the next digit of the values that share the rank's prefix found so
// far; a one-block-per-item kernel then picks the digit holding the rank.  The
// result is the exact order statistic, the same value select_nth_unstable_by
// returns in the reference, independent of any ordering of the map.  The ranks
// themselves are computed on the device from the item's count, so the whole
// selection runs without the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "bm_device.h"
#include "bm_engine.h"

namespace bm {

// The quantiles in the order of src/metrics.rs:242-255.
__constant__ double c_qs[5] = {0.5, 0.1, 0.25, 0.75, 0.9};
static const double h_qs[5] = {0.5, 0.1, 0.25, 0.75, 0.9};

// Ranks of linear_quantile_unstable (src/metrics.rs:271-291): slot 2q holds
// floor(q (count - 1)), slot 2q + 1 ceil(q (count - 1)).  One block per item.
__global__ void __launch_bounds__(256) k_select_init(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[blockIdx.x];
  if (!(I.flags & kItemQuantiles)) return;
  BM_GLOBAL SelectState* __restrict__ st = gbl(I.sel);
  const unsigned long long count = gbl(I.stats)->count;
  if (threadIdx.x < 5) {
    const uint32_t q = threadIdx.x;
    const double position = c_qs[q] * (double)(count > 0 ? count - 1 : 0);
    st->rank[2 * q] = (unsigned long long)floor(position);
    st->rank[2 * q + 1] = (unsigned long long)ceil(position);
    st->prefix[2 * q] = 0;
    st->prefix[2 * q + 1] = 0;
  }
  if (threadIdx.x == 0) {
    st->nranks = count > 0 ? (uint32_t)kMaxRanks : 0u;
    st->pad = 0;
  }
}

// Inclusive scan of one u64 per thread over a 256-thread block.
__device__ __forceinline__ unsigned long long block_scan256(unsigned long long v,
                                                            unsigned long long* s_w) {
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(v, off);
    if (lane >= (uint32_t)off) v += y;
  }
  __syncthreads();
  if (lane == 63) s_w[wv] = v;
  __syncthreads();
  for (uint32_t i = 0; i < wv; ++i) v += s_w[i];
  return v;
}

// First digit from the clean-up's 16-bit histogram: bin b counts values whose
// top 16 bits are kTopBase + b.  Each thread owns 8 consecutive bins.  One
// block per item; also clears the digit histograms.
__global__ void __launch_bounds__(256)
k_top16_pick(const ItemDesc* __restrict__ items) {
  __shared__ unsigned long long s_w[4];
  const ItemDesc& I = items[blockIdx.x];
  if (!(I.flags & kItemQuantiles)) return;
  BM_GLOBAL SelectState* __restrict__ st = gbl(I.sel);
  BM_GLOBAL const uint32_t* __restrict__ top16 = gbl(I.top16);
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) gbl(I.qhist)[i] = 0;
  constexpr int per = kTopBins / 256;
  uint32_t c[per];
  unsigned long long own = 0;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    c[k] = top16[threadIdx.x * per + k];
    own += c[k];
  }
  const unsigned long long incl = block_scan256(own, s_w);
  const unsigned long long excl = incl - own;
  const uint32_t nr = st->nranks;
  __syncthreads();
  for (uint32_t r = 0; r < nr; ++r) {
    const unsigned long long k = st->rank[r];
    __syncthreads();
    if (k >= excl && k < incl) {  // exactly one thread holds the rank
      unsigned long long cum = excl;
      uint32_t b = 0;
      for (; b < (uint32_t)per - 1; ++b) {
        if (k < cum + c[b]) break;
        cum += c[b];
      }
      st->prefix[r] = (threadIdx.x * per + b + kTopBase) << 16;
      st->rank[r] = k - cum;
    }
  }
}

// One 8-bit digit (bits shift..shift+7) of the values sharing each rank's
// prefix; ranks with equal prefixes share one histogram row.
//
// A value can share a prefix only if its top 16 bits name the top-16 bin of
// one (a bin of the clean-up's histogram).  A table over those bins gives the
// bin's group and the group's prefix bits below the top 16, so a value is
// classified with one LDS read at an index clamped into the table, whose
// extra last entry (no group) takes every value outside the histogram's range
// (zero, beyond it, infinite: their top 16 bits minus kTopBase wrap or exceed
// kTopBins); the reads of a trip's 16 values have no branch between them.  A bin holding two
// groups' prefixes (they then differ below bit 16) is flagged and its values
// search the groups.
//
// Block b of the first pass (shift 8) reads the map's range [b qper, (b + 1)
// qper) and keeps the values that share a prefix (a few per cent of the map)
// at the same offsets of cand, with their count in qcnt[b]; block b of the
// second pass reads only those.  Every value sharing a prefix of the second
// pass shares the first pass's prefix of its rank, so the second pass sees
// all of them.
constexpr uint32_t kNoGroup = 0xFFu, kMultiGroup = 0x100u;
__global__ void __launch_bounds__(256)
k_digit_hist(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
             uint32_t nitems, int shift) {
  __shared__ uint32_t h[kMaxRanks][256];
  __shared__ uint32_t gpre[kMaxRanks], grow[kMaxRanks], s_ng, s_kept, s_multi;
  // per top-16 bin: group (kNoGroup: none) | kMultiGroup | prefix bits 0..15 << 16
  __shared__ uint32_t s_btab[kTopBins + 1];
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const SelectState* __restrict__ st = gbl(I.sel);
  const bool first = shift == 8;
  BM_GLOBAL uint32_t* __restrict__ cand = gbl(I.cand);
  BM_GLOBAL unsigned long long* __restrict__ hist = gbl(I.qhist);
  const uint32_t nr = st->nranks;
  if (nr == 0) return;
  const uint64_t lo = (uint64_t)vb * I.qper;
  const uint64_t hi = first ? min(I.n, lo + I.qper) : lo + gbl(I.qcnt)[vb];
  BM_GLOBAL const uint32_t* __restrict__ src =
      first ? (BM_GLOBAL const uint32_t*)gbl(I.map) : (BM_GLOBAL const uint32_t*)cand;
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  for (uint32_t i = threadIdx.x; i <= (uint32_t)kTopBins; i += blockDim.x) s_btab[i] = kNoGroup;
  const uint32_t hi_mask = shift >= 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
  __syncthreads();
  if (threadIdx.x == 0) {  // distinct prefixes, and the table of their top-16 bins
    uint32_t ng = 0, multi = 0;
    for (uint32_t r = 0; r < nr; ++r) {
      const uint32_t pre = st->prefix[r];
      uint32_t g = 0;
      while (g < ng && gpre[g] != pre) ++g;
      grow[r] = g;
      if (g < ng) continue;
      gpre[ng++] = pre;
      // (a rank k_top16_pick did not place -- values beyond the histogram, a
      // case the statistics report as an error -- keeps prefix 0: no bin)
      const uint32_t bin = (pre >> 16) - kTopBase;
      if (bin >= (uint32_t)kTopBins) continue;
      const uint32_t e = s_btab[bin];
      s_btab[bin] = (e & 0xFFu) == kNoGroup ? (g | (pre << 16)) : (e | kMultiGroup);
      multi |= (e & 0xFFu) != kNoGroup;
    }
    s_ng = ng;
    s_multi = multi;
    s_kept = 0;
  }
  __syncthreads();
  const uint32_t ng = s_ng;
  const bool any_multi = s_multi != 0;  // a bin holds two groups' prefixes (rare)
  const uint32_t lo_mask = hi_mask & 0xFFFFu;  // prefix bits below the top 16
  const uint32_t lane = threadIdx.x & 63;
  // four consecutive values per 16-byte load (the map and cand are 256-byte
  // aligned and qper is a multiple of 4096), and kLoads such loads per thread
  // and trip, each coalesced across the wave and all issued before the first
  // is used
  constexpr int kLoads = 4, kTrip = kLoads * 256 * 4;
  for (uint64_t base = lo; base < hi; base += kTrip) {
    uint32_t b[kLoads][4];
    const uint64_t p0 = base + 4 * threadIdx.x;
    if (base + kTrip <= hi) {  // the whole trip in range (uniform)
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const v4u32 w = *(BM_GLOBAL const v4u32*)(src + p0 + 1024 * u);
        b[u][0] = w.x;
        b[u][1] = w.y;
        b[u][2] = w.z;
        b[u][3] = w.w;
      }
    } else {
#pragma unroll
      for (int u = 0; u < kLoads; ++u)
        for (int k = 0; k < 4; ++k) {
          const uint64_t i = p0 + 1024 * u + k;
          b[u][k] = i < hi ? src[i] : 0u;
        }
    }
    uint32_t t[kLoads][4];
#pragma unroll
    for (int u = 0; u < kLoads; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        t[u][k] = s_btab[min((b[u][k] >> 16) - kTopBase, (uint32_t)kTopBins)];
    uint32_t keep = 0;  // bit 4u + k: value (u, k) shares a prefix
#pragma unroll
    for (int u = 0; u < kLoads; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t v = b[u][k];
        // (values outside the table's range read its last entry: no group)
        const uint32_t e = t[u][k];
        uint32_t g = (v & lo_mask) == (e >> 16) ? (e & 0xFFu) : kNoGroup;
        const bool multi = e & kMultiGroup;
        if (any_multi && __ballot(multi)) {
          if (multi) {
            g = kNoGroup;
            for (uint32_t q = 0; q < ng; ++q)
              if ((v & hi_mask) == gpre[q]) g = q;
          }
        }
        // few lanes hold a candidate, so a plain LDS atomic per candidate
        if (g < ng) {
          atomicAdd(&h[g][(v >> shift) & 0xFFu], 1u);
          keep |= 1u << (4 * u + k);
        }
      }
    if (first) {
      // this wave's kept values go to consecutive offsets from one LDS
      // reservation: the lane's offset is the sum of the kept counts of the
      // lanes below it, bit by bit of the count (at most 16)
      const uint32_t c = __popc(keep);
      uint32_t below = 0, total = 0;
#pragma unroll
      for (int bit = 0; bit < 5; ++bit) {
        const unsigned long long m = __ballot((c >> bit) & 1u);
        const uint32_t lower =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        below += lower << bit;
        total += (uint32_t)__popcll(m) << bit;
      }
      if (total) {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&s_kept, total);
        at = __shfl(at, 0) + below;
        BM_GLOBAL uint32_t* __restrict__ dst = cand + lo;
#pragma unroll
        for (int u = 0; u < kLoads; ++u)
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (keep & (1u << (4 * u + k))) dst[at++] = b[u][k];
      }
    }
  }
  __syncthreads();
  if (first && threadIdx.x == 0) gbl(I.qcnt)[vb] = s_kept;
  for (uint32_t i = threadIdx.x; i < ng * 256; i += blockDim.x) {
    const uint32_t v = (&h[0][0])[i];
    if (!v) continue;
    // expand the shared row to every rank of the group
    const uint32_t gi = i / 256, d = i % 256;
    for (uint32_t r = 0; r < nr; ++r)
      if (grow[r] == gi) atomicAdd((unsigned long long*)&hist[r * 256 + d], (unsigned long long)v);
  }
}

// The digit holding each rank; one block (256 threads = 256 digits) per item.
// Clears the histograms for the next pass.
__global__ void __launch_bounds__(256) k_digit_pick(const ItemDesc* __restrict__ items, int shift) {
  __shared__ unsigned long long s_w[4];
  const ItemDesc& I = items[blockIdx.x];
  if (!(I.flags & kItemQuantiles)) return;
  BM_GLOBAL SelectState* __restrict__ st = gbl(I.sel);
  BM_GLOBAL unsigned long long* __restrict__ hist = gbl(I.qhist);
  const uint32_t nr = st->nranks;
  const uint32_t d = threadIdx.x;
  for (uint32_t r = 0; r < nr; ++r) {
    const unsigned long long c = hist[r * 256 + d];
    const unsigned long long incl = block_scan256(c, s_w);
    const unsigned long long excl = incl - c;
    const unsigned long long k = st->rank[r];
    __syncthreads();
    if (k >= excl && k < incl) {
      st->prefix[r] |= d << shift;
      st->rank[r] = k - excl;
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) hist[i] = 0;
}

// linear_quantile_unstable (src/metrics.rs:271-291) on the selected values.
static double linear_quantile(double q, uint64_t len, double lower_value, double upper_value,
                              float scale) {
  const double position = q * (double)(len - 1);
  const uint64_t lower = (uint64_t)std::floor(position);
  const double fraction = position - (double)lower;
  const double quantile_value = lower_value + fraction * (upper_value - lower_value);
  return (double)((float)quantile_value * scale);
}

QuantileResult quantiles_from(const SelectState& sel, uint64_t count, float scale_f32) {
  QuantileResult r;
  if (count == 0 || sel.nranks != (uint32_t)kMaxRanks) return r;
  double v[kMaxRanks];
  for (int i = 0; i < kMaxRanks; ++i) {
    float f;
    std::memcpy(&f, &sel.prefix[i], 4);
    v[i] = (double)f;
  }
  double* out[5] = {&r.median, &r.p10, &r.p25, &r.p75, &r.p90};
  for (int q = 0; q < 5; ++q)
    *out[q] = linear_quantile(h_qs[q], count, v[2 * q], v[2 * q + 1], scale_f32);
  return r;
}

}  // namespace bm
