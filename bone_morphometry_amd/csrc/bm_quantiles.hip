// bm_quantiles.hip -- radix select of up to 10 order statistics at once.
//
// The values are positive finite f32, whose bit patterns sort like the values,
// so the k-th smallest value is found digit by digit (4 passes of 8 bits, most
// significant first).  Each pass histograms, for every wanted rank, the next
// digit of the values that share the rank's prefix found so far; a one-block
// kernel then picks the digit holding the rank.  The result is the exact
// order statistic, the same value select_nth_unstable_by returns in the
// reference (src/metrics.rs:293-298), independent of any ordering of the map.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bm_device.h"
#include "bm_quantiles.h"

namespace bm {

constexpr int kMaxRanks = 10;

struct SelectState {
  uint32_t prefix[kMaxRanks];
  unsigned long long rank[kMaxRanks];  // rank within the values sharing prefix
  uint32_t nranks;
  uint32_t pad;
};

// One 8-bit digit (bits shift..shift+7) of the values sharing each rank's
// prefix; ranks with equal prefixes share one histogram row.
__global__ void __launch_bounds__(256)
k_digit_hist(const float* __restrict__ map, uint64_t n, const SelectState* __restrict__ st,
             int shift, unsigned long long* __restrict__ hist /* [kMaxRanks][256] */) {
  __shared__ uint32_t h[kMaxRanks][256];
  __shared__ uint32_t gpre[kMaxRanks], grow[kMaxRanks], s_ng;
  const uint32_t nr = st->nranks;
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  const uint32_t hi_mask = shift >= 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
  if (threadIdx.x == 0) {  // distinct prefixes
    uint32_t ng = 0;
    for (uint32_t r = 0; r < nr; ++r) {
      uint32_t g = 0;
      while (g < ng && gpre[g] != st->prefix[r]) ++g;
      if (g == ng) gpre[ng++] = st->prefix[r];
      grow[r] = g;
    }
    s_ng = ng;
  }
  __syncthreads();
  const uint32_t ng = s_ng;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t iters = (n + stride - 1) / stride;  // uniform trip count (wave ballots)
  for (uint64_t it = 0; it < iters; ++it) {
    const uint64_t i = it * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t g = ng, digit = 0;
    if (i < n) {
      const uint32_t b = __float_as_uint(map[i]);
      if (b != 0 && b < 0x7F800000u) {  // positive finite values only
        digit = (b >> shift) & 0xFFu;
        for (uint32_t k = 0; k < ng; ++k)
          if ((b & hi_mask) == gpre[k]) g = k;
      }
    }
    wave_hist_add(&h[0][0], g < ng, g * 256 + digit);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ng * 256; i += blockDim.x) {
    const uint32_t v = (&h[0][0])[i];
    if (!v) continue;
    // expand the shared row to every rank of the group
    const uint32_t gi = i / 256, d = i % 256;
    for (uint32_t r = 0; r < nr; ++r)
      if (grow[r] == gi) atomicAdd(&hist[r * 256 + d], (unsigned long long)v);
  }
}

__global__ void k_digit_pick(SelectState* __restrict__ st, int shift,
                             unsigned long long* __restrict__ hist) {
  const uint32_t r = threadIdx.x;
  if (r < st->nranks) {
    unsigned long long k = st->rank[r], cum = 0;
    uint32_t d = 0;
    for (; d < 256; ++d) {
      const unsigned long long c = hist[r * 256 + d];
      if (k < cum + c) break;
      cum += c;
    }
    st->prefix[r] |= d << shift;
    st->rank[r] = k - cum;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) hist[i] = 0;
}

// First level from the 16-bit histogram the clean-up kernel accumulates:
// bin b counts values whose top 16 bits are kTopBase + b.
__global__ void k_top16_pick(SelectState* __restrict__ st, const uint32_t* __restrict__ top16,
                             unsigned long long* __restrict__ hist) {
  const uint32_t r = threadIdx.x;
  if (r < st->nranks) {
    unsigned long long k = st->rank[r], cum = 0;
    uint32_t d = 0;
    for (; d < (uint32_t)kTopBins; ++d) {
      const unsigned long long c = top16[d];
      if (k < cum + c) break;
      cum += c;
    }
    st->prefix[r] = (d + kTopBase) << 16;
    st->rank[r] = k - cum;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) hist[i] = 0;
}

__global__ void k_select_init(SelectState* __restrict__ st, unsigned long long* __restrict__ hist,
                              SelectState init) {
  if (threadIdx.x == 0) *st = init;
  for (uint32_t i = threadIdx.x; i < kMaxRanks * 256; i += blockDim.x) hist[i] = 0;
}

// linear_quantile_unstable (src/metrics.rs:271-291) on selected order statistics.
static double linear_quantile(double q, uint64_t len, const double* vals, int lo_slot,
                              int hi_slot, float scale) {
  const double position = q * (double)(len - 1);
  const uint64_t lower = (uint64_t)std::floor(position);
  const double fraction = position - (double)lower;
  const double lower_value = vals[lo_slot];
  const double upper_value = vals[hi_slot];
  const double quantile_value = lower_value + fraction * (upper_value - lower_value);
  return (double)((float)quantile_value * scale);
}

QuantileResult device_quantiles(Engine& e, const float* d_map, uint64_t n, uint64_t count,
                                float scale_f32, hipStream_t s, const uint32_t* top16) {
  static const double qs[5] = {0.5, 0.1, 0.25, 0.75, 0.9};
  SelectState init{};
  int lo_slot[5], hi_slot[5];
  auto slot_of = [&](unsigned long long rank) {
    for (uint32_t i = 0; i < init.nranks; ++i)
      if (init.rank[i] == rank) return (int)i;
    init.rank[init.nranks] = rank;
    init.prefix[init.nranks] = 0;
    return (int)init.nranks++;
  };
  for (int i = 0; i < 5; ++i) {
    const double position = qs[i] * (double)(count - 1);
    const uint64_t lower = (uint64_t)std::floor(position);
    const uint64_t upper = (uint64_t)std::ceil(position);
    lo_slot[i] = slot_of(lower);
    hi_slot[i] = upper == lower ? lo_slot[i] : slot_of(upper);
  }
  char* scratch = static_cast<char*>(
      e.scratch_bytes(sizeof(SelectState) + kMaxRanks * 256 * sizeof(unsigned long long) + 256, 2));
  SelectState* d_st = reinterpret_cast<SelectState*>(scratch);
  unsigned long long* d_hist = reinterpret_cast<unsigned long long*>(scratch + 256);
  k_select_init<<<1, 256, 0, s>>>(d_st, d_hist, init);
  unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
  int first = 24;
  if (top16) {  // the top 16 bits are already counted: two 8-bit passes remain
    BM_LAUNCH("quantile_pick", k_top16_pick, 1, 256, 0, s, d_st, top16, d_hist);
    first = 8;
  }
  for (int shift = first; shift >= 0; shift -= 8) {
    BM_LAUNCH("quantile_hist", k_digit_hist, grid, 256, 0, s, d_map, n, d_st, shift, d_hist);
    BM_LAUNCH("quantile_pick", k_digit_pick, 1, 256, 0, s, d_st, shift, d_hist);
  }
  SelectState out{};
  BM_HIP(hipMemcpyAsync(&out, d_st, sizeof out, hipMemcpyDeviceToHost, s));
  BM_HIP(hipStreamSynchronize(s));
  double vals[kMaxRanks];
  for (uint32_t i = 0; i < out.nranks; ++i) {
    float f;
    std::memcpy(&f, &out.prefix[i], 4);
    vals[i] = (double)f;
  }
  QuantileResult r;
  r.median = linear_quantile(qs[0], count, vals, lo_slot[0], hi_slot[0], scale_f32);
  r.p10 = linear_quantile(qs[1], count, vals, lo_slot[1], hi_slot[1], scale_f32);
  r.p25 = linear_quantile(qs[2], count, vals, lo_slot[2], hi_slot[2], scale_f32);
  r.p75 = linear_quantile(qs[3], count, vals, lo_slot[3], hi_slot[3], scale_f32);
  r.p90 = linear_quantile(qs[4], count, vals, lo_slot[4], hi_slot[4], scale_f32);
  return r;
}

}  // namespace bm
