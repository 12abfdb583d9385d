// bm_tiles.hip -- tile kernels of local thickness: ridges, sphere fill, clean-up.
//
// The volume is cut into 8x8x8 tiles.  The ridge and clean-up passes stream
// 64 x 8 (x, y) footprints along z (rolling windows of planes in LDS); the
// sphere fill works per tile (one 512-thread workgroup, one wavefront per
// 4x4x4 sub-brick).
//
// Ridges (src/local_thickness.rs:104-167).  Each voxel's 26 neighbours are
// tested against the templates of its squared distance.  Each tile's ridges
// are written as a list sorted by decreasing r^2 (counting sort on r^2 >>
// shift), entry = r^2 << 9 | local index, at a fixed slot of 512 entries per
// tile.  Two bit masks per tile are written too: the selected voxels (d2 > 0)
// and the fill targets (selected voxels and non-selected voxels with a
// selected face neighbour).
//
// Sphere fill (src/local_thickness.rs:169-208).  M[v] = max r^2(c) over ridges
// c with |v - c|^2 <= r^2(c).  Instead of scattering every ball, each tile
// gathers the ridges whose ball meets it from the sorted lists of the tiles in
// reach, orders them by decreasing r^2 in LDS, and every 4x4x4 sub-brick scans
// them until all of its targets are covered: in decreasing order the first
// (bucket of) covering ball(s) gives the maximum.  Exactness of the exit test:
//   * a selected voxel v is covered by its own ball, hence by a ridge ball
//     (the ridge filter only drops a ball contained in a neighbour's larger
//     ball), so it is covered by some candidate;
//   * a non-selected voxel s is touched by a ball of centre c only if s is a
//     nearest site of c; if s had no selected face neighbour, every in-bounds
//     face neighbour s+e would be a site and |s+e-c|^2 < |s-c|^2 whenever
//     (c-s).e > 0, so c - s would have no positive component in any direction,
//     i.e. c = s -- impossible.  If s has a selected face neighbour u, then
//     d2(u) = 1 and the unit ball of u contains s, so M[s] >= 1.
// Hence targets are exactly the voxels with M > 0, every target is covered by
// some candidate, and all other voxels have M = 0.  Balls of radius above
// kGatherMaxRadius are scattered after the gather (k_fill_scatter_big).
//
// Clean-up (src/local_thickness.rs:210-281) + phase mask (:31-43) + the value
// part of nonzero_stats (src/metrics.rs:215-226): surface marks from the zero
// voxels of three planes, then the f32 neighbour means in NEIGHBORS_26 order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bm_device.h"
#include "bm_engine.h"
#include "bm_tiles.h"

namespace bm {

__device__ __forceinline__ uint32_t bitlen(uint32_t v) { return v ? 32 - __clz(v) : 0; }

// Position -> tile mapping of an item in the given tile order, with the
// item's divisor constants (no runtime division).
__device__ __forceinline__ TileWalk walk_of(const ItemDesc& I, bool grouped) {
  const uint32_t* a = I.divs[grouped ? 2 : 0];
  const uint32_t* b = I.divs[grouped ? 3 : 1];
  return TileWalk(I.TX, I.TY, I.TZ, grouped, FastDiv(a[0], a[1], a[2], a[3]),
                  FastDiv(b[0], b[1], b[2], b[3]));
}

// Shift so that (max_r2 >> shift) < kBuckets.
__device__ __forceinline__ uint32_t bucket_shift(uint32_t max_r2) {
  const uint32_t b = bitlen(max_r2);
  return b > kBucketBits ? b - kBucketBits : 0;
}

// ---------------------------------------------------------------------------
// Ridges streamed along z (src/local_thickness.rs:104-167).  A work item is a 64 x 8 (x, y) footprint --
// eight tiles along x -- over kRidgeZ planes (whole tile layers).  Planes of
// d^2 with a one-voxel halo pass through a ring of four LDS slots: while
// plane s is evaluated against planes s-1 and s+1, plane s+2 is in flight and
// plane s+1's successor takes the free slot, so one barrier per plane.  A
// ridge voxel's entry (r^2 << 9 | local index) is appended to its tile's LDS
// list (one LDS atomic per eight-lane group) and counted in the tile's bucket
// histogram; when a tile layer is complete, wavefront k sorts tile k's
// entries by bucket into the tile's list slot and writes its header, and the
// tile masks of each plane are written once the plane is done.
constexpr int kTmplLds = 1024;
__global__ void __launch_bounds__(512)
k_ridges_stream(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                uint32_t nitems) {
  constexpr int kRX = 66, kRY = 10, kR = kRX * kRY;
  __shared__ uint32_t s_d2[4][kR];
  __shared__ uint32_t s_ent[8][512];
  __shared__ uint32_t s_cnt[8][kBuckets];
  __shared__ unsigned long long s_msk[2][8][2];  // plane parity, tile, (selected, target)
  __shared__ uint32_t s_n[8], s_mx[8];
  __shared__ uint16_t s_pref[8][16];
  // the three templates of r^2 < kTmplLds as 16-bit fields (t1 | t2 << 16, t3)
  __shared__ uint2 s_tm[kTmplLds];
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const uint32_t* __restrict__ d2 = gbl(I.d2);
  const uint32_t W = I.W, H = I.H, D = I.D, TX = I.TX, TY = I.TY, bound = I.bound;
  BM_GLOBAL const uint32_t* __restrict__ tmpl = gbl(I.tmpl);
  BM_GLOBAL PhaseStatus* __restrict__ st = gbl(I.status);
  BM_GLOBAL uint32_t* __restrict__ lists = gbl(I.lists);
  BM_GLOBAL TileMeta* __restrict__ meta = gbl(I.meta);
  BM_GLOBAL unsigned long long* __restrict__ masks = gbl(I.masks);
  if (!st->has_site) return;
  const uint32_t nfx = (W + 63) / 64, nfy = (H + 7) / 8;
  const uint32_t npos = ridge_work_items(W, H, D);
  const uint32_t j = xcd_tile(vb, npos);
  if (j >= npos) return;
  // neighbouring footprints along y run on one XCD at once (shared halos)
  const uint32_t fy = j % nfy, fx = (j / nfy) % nfx, zc = j / (nfy * nfx);
  const int x0 = (int)fx * 64, y0 = (int)fy * 8, za = (int)(zc * kRidgeZ),
            zb = min(za + (int)kRidgeZ, (int)D);
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const bool valid = x0 + (int)lane < (int)W && y0 + (int)wv < (int)H;
  const uint32_t shift = bucket_shift(st->max_d2);  // one key for every tile
  const uint32_t top = st->max_d2 >> shift;
  // templates of the phase's radii in LDS when they fit (every template is
  // then below 2^16: T(m) <= m + 2 sqrt(3 m) + 3 for m < kTmplLds)
  const bool lds_tmpl = st->max_d2 < (uint32_t)kTmplLds;
  if (lds_tmpl)
    for (uint32_t i = threadIdx.x; i <= st->max_d2; i += blockDim.x)
      s_tm[i] = make_uint2(tmpl[i] | (tmpl[(bound + 1) + i] << 16), tmpl[2 * (bound + 1) + i]);
  for (uint32_t i = t; i < 8u * kBuckets; i += blockDim.x) (&s_cnt[0][0])[i] = 0;
  if (t < 8) {
    s_n[t] = 0;
    s_mx[t] = 0;
  }
  const bool has1 = t + 512 < (uint32_t)kR;
  // (the empty asm keeps the element coordinates from being hoisted and held)
  auto load = [&](uint32_t e, int z) -> uint32_t {
    asm volatile("" : "+v"(e));
    const int qx = x0 - 1 + (int)(e % kRX), qy = y0 - 1 + (int)(e / kRX);
    // outside the array: never >= a template, never selected
    if (z < 0 || z >= (int)D || qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) return 0u;
    return d2[((size_t)z * H + qy) * W + qx];
  };
  s_d2[(za - 1) & 3][t] = load(t, za - 1);
  if (has1) s_d2[(za - 1) & 3][t + 512] = load(t + 512, za - 1);
  s_d2[za & 3][t] = load(t, za);
  if (has1) s_d2[za & 3][t + 512] = load(t + 512, za);
  uint32_t pre0 = load(t, za + 1), pre1 = has1 ? load(t + 512, za + 1) : 0u;
  const uint32_t ci = (wv + 1) * kRX + lane + 1;  // this voxel in a plane slot
  const uint32_t k = lane >> 3;                   // its tile in the footprint
  auto write_masks = [&](int z) {                 // plane z's tile mask words
    if (t < 16) {
      const uint32_t kk = t >> 1, which = t & 1, tx = fx * 8 + kk;
      if (tx < TX)
        masks[((size_t)((uint32_t)z / 8 * TY + fy) * TX + tx) * 16 + which * 8 + ((uint32_t)z & 7)] =
            s_msk[z & 1][kk][which];
    }
  };
  for (int s = za; s < zb; ++s) {
    s_d2[(s + 1) & 3][t] = pre0;
    if (has1) s_d2[(s + 1) & 3][t + 512] = pre1;
    if (s + 2 <= zb) {  // the plane after next, in flight meanwhile
      pre0 = load(t, s + 2);
      pre1 = has1 ? load(t + 512, s + 2) : 0u;
    }
    __syncthreads();
    if (s > za) write_masks(s - 1);
    const uint32_t* pm = s_d2[(s - 1) & 3];
    const uint32_t* pc = s_d2[s & 3];
    const uint32_t* pp = s_d2[(s + 1) & 3];
    const uint32_t r2 = pc[ci];  // 0 outside the array
    const bool sel = r2 != 0;
    // the neighbours and templates are read by every lane, selected or not
    // (r2 = 0 reads the tables' first entries), so no branch separates reads
    uint32_t t1, t2, t3;
    if (lds_tmpl) {
      const uint2 tt = s_tm[r2];
      t1 = tt.x & 0xFFFFu;
      t2 = tt.x >> 16;
      t3 = tt.y;
    } else {
      t1 = tmpl[r2];
      t2 = tmpl[(bound + 1) + r2];
      t3 = tmpl[2 * (bound + 1) + r2];
    }
    uint32_t nb[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
      const uint32_t* pl = dz < 0 ? pm : dz > 0 ? pp : pc;
      nb[q] = q == 13 ? 0u : pl[(int)ci + dx + kRX * dy];
    }
    // fill target: selected, or in the image with a selected face neighbour
    const bool target =
        sel || (valid && (nb[4] | nb[10] | nb[12] | nb[14] | nb[16] | nb[22]) != 0);
    // A ridge unless a neighbour's squared distance reaches the template of
    // its offset class (1: face, 2: edge, 3: corner neighbour).
    uint32_t mx[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const int kc = (q % 3 != 1) + ((q / 3) % 3 != 1) + (q / 9 != 1);
      mx[kc] = max(mx[kc], nb[q]);
    }
    const bool ridge = sel && !(mx[1] >= t1 || mx[2] >= t2 || mx[3] >= t3);
    // mask bytes: row wv of each of the eight tiles (x + 8y within a plane)
    const unsigned long long msel = __ballot(sel), mtgt = __ballot(target);
    if (lane < 8) {
      reinterpret_cast<uint8_t*>(&s_msk[s & 1][lane][0])[wv] = (uint8_t)(msel >> (8 * lane));
      reinterpret_cast<uint8_t*>(&s_msk[s & 1][lane][1])[wv] = (uint8_t)(mtgt >> (8 * lane));
    }
    // the ridge's entry, appended to its tile's list
    const unsigned long long rb = __ballot(ridge);
    const uint32_t grp = (uint32_t)(rb >> (8 * k)) & 0xFFu;
    uint32_t base = 0;
    if ((lane & 7) == 0 && grp) base = atomicAdd(&s_n[k], (uint32_t)__popc(grp));
    base = (uint32_t)__shfl((int)base, (int)(lane & ~7u));
    if (ridge) {
      const uint32_t li = (lane & 7) | (wv << 3) | (((uint32_t)s & 7) << 6);
      s_ent[k][base + __popc(grp & ((1u << (lane & 7)) - 1u))] = (r2 << 9) | li;
      atomicAdd(&s_cnt[k][top - (r2 >> shift)], 1u);
      atomicMax(&s_mx[k], r2);
    }
    if ((s & 7) != 7 && s != zb - 1) continue;
    // Tile layer complete: wavefront wv sorts tile wv.
    __syncthreads();
    {
      const uint32_t kk = wv, tx = fx * 8 + kk;
      if (tx < TX) {
        const uint32_t tile = ((uint32_t)s / 8 * TY + fy) * TX + tx;
        uint32_t* cnt = s_cnt[kk];
        // exclusive scan of the counts, 8 buckets per lane
        static_assert(kBuckets == 64 * 8, "the tile-layer scan holds 8 buckets per lane");
        uint32_t local[8], sum = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          local[i] = cnt[lane * 8 + i];
          sum += local[i];
        }
        const uint32_t x = wave_inclusive_sum(sum);
        uint32_t run = x - sum;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          cnt[lane * 8 + i] = run;
          run += local[i];
        }
        const uint32_t count = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // prefix counts: entries with key >= kPrefR2(j) >> shift are the
        // buckets 0 .. top - (kPrefR2(j) >> shift): the next bucket's start
        if (lane < 16) {
          const uint32_t kj = kPrefR2((int)lane) >> shift;
          uint32_t p = 0;
          if (kj <= top) {
            const uint32_t nbk = top - kj + 1;
            p = nbk < (uint32_t)kBuckets ? cnt[nbk] : count;
          }
          s_pref[kk][lane] = (uint16_t)p;
        }
        const uint32_t mxr = s_mx[kk];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) {
          TileMeta m;
          m.count = count;
          m.maxr2 = mxr;
          for (int q = 0; q < 16; ++q) m.pref[q] = s_pref[kk][q];
          *(TileMeta*)&meta[tile] = m;
          if (mxr >= I.r2_limit) atomicOr((uint32_t*)&st->error, 1u);
          if (mxr > kGatherMaxR2) {  // balls the gather leaves to the scatter
            BM_GLOBAL uint32_t* __restrict__ big = gbl(I.bigtiles);
            big[1 + atomicAdd((uint32_t*)&big[0], 1u)] = tile;
          }
        }
        // the entries at their sorted positions (any order inside a bucket)
        for (uint32_t e = lane; e < count; e += 64) {
          const uint32_t ent = s_ent[kk][e];
          const uint32_t pos = atomicAdd(&cnt[top - ((ent >> 9) >> shift)], 1u);
          lists[(size_t)tile * 512 + pos] = ent;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      // cleared for the next layer
      for (uint32_t i = lane; i < (uint32_t)kBuckets; i += 64) s_cnt[kk][i] = 0;
      if (lane == 0) {
        s_n[kk] = 0;
        s_mx[kk] = 0;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  write_masks(zb - 1);
  // planes of the last tile layer beyond the array: empty mask words
  if (zb == (int)D && t < 16) {
    const uint32_t kk = t >> 1, which = t & 1, tx = fx * 8 + kk;
    if (tx < TX)
      for (uint32_t z = D; z < (D + 7) / 8 * 8; ++z)
        masks[((size_t)(z / 8 * TY + fy) * TX + tx) * 16 + which * 8 + (z & 7)] = 0ull;
  }
}

// ---------------------------------------------------------------------------
// A candidate's packed position: its offset from the block origin plus 512 in
// bits 0-9 (x), 10-19 (y) and 20-29 (z).
// Squared length of a gap vector.  Gaps are below 2^15 (axes are at most 16384
// voxels), so 24-bit multiplies are exact and run at the full vector rate.
__device__ __forceinline__ uint32_t sq24(uint32_t v) {
  // (written out: the compiler turns __umul24 of an operand it can bound
  // into a full 32-bit multiply, a quarter-rate instruction)
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %1" : "=v"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ uint32_t gap2(uint32_t ex, uint32_t ey, uint32_t ez) {
  return sq24(ex) + sq24(ey) + sq24(ez);
}

// Number of entries of a tile list that can reach a box at squared distance
// bd2: those whose bucket key is at least that of kPrefR2(j), j the largest
// with kPrefR2(j) <= bd2 (an upper bound read from pref[]).
__device__ __forceinline__ uint32_t reach_len(BM_GLOBAL const TileMeta* __restrict__ m, uint32_t count,
                                              uint32_t bd2) {
  if (bd2 < 1) return count;
  // (2j+1)^2 <= bd2  <=>  j <= (isqrt(bd2) - 1) / 2; bd2 < 3 * 2^28 (gaps of
  // at most 16384), so the root is below 2^15 and its checks fit 32 bits
  uint32_t r = (uint32_t)sqrtf((float)bd2);
  while (__umul24(r + 1, r + 1) <= bd2) ++r;
  while (__umul24(r, r) > bd2) --r;
  const uint32_t j = min((r - 1) >> 1, 15u);
  return m->pref[j];
}

// Exclusive scan by one wavefront of val(0 .. n-1), n <= 64 * kPer, kPer
// consecutive values per lane: out(j, prefix) receives each exclusive prefix;
// returns the total (in every lane).  No barrier: the other wavefronts wait
// at the caller's next one.
template <int kPer, typename Val, typename Out>
__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t n, uint32_t lane, Val val, Out out) {
  uint32_t local[kPer], sum = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const uint32_t j = lane * kPer + (uint32_t)i;
    local[i] = j < n ? val(j) : 0u;
    sum += local[i];
  }
  const uint32_t x = wave_inclusive_sum(sum);
  uint32_t run = x - sum;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const uint32_t j = lane * kPer + (uint32_t)i;
    if (j < n) out(j, run);
    run += local[i];
  }
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// The same scan with the fewest values per lane that hold n (<= 64 kMax).
template <int kMax, typename Val, typename Out>
__device__ __forceinline__ uint32_t wave_scan_sized(uint32_t n, uint32_t lane, Val val, Out out) {
  if (n <= 64) return wave_exclusive_scan<1>(n, lane, val, out);
  if (kMax >= 2 && n <= 128) return wave_exclusive_scan<(kMax >= 2 ? 2 : 1)>(n, lane, val, out);
  if (kMax >= 4 && n <= 256) return wave_exclusive_scan<(kMax >= 4 ? 4 : 1)>(n, lane, val, out);
  return wave_exclusive_scan<kMax>(n, lane, val, out);
}

// One block per TXN x-adjacent tiles (a 8*TXN x 8 x 8 target box, 512*TXN
// threads, one wavefront per 4x4x4 sub-brick): the tiles share the gathering
// and sorting of their candidates.  kDiag: work counters and phase clocks into
// the item's diag buffer (a separate instantiation, so the counters cost
// nothing otherwise).
template <bool kDiag, int TXN>
__device__ __forceinline__ void
k_fill_gather_t(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                uint32_t nitems) {
  constexpr int NW = 8 * TXN;                // wavefronts
  constexpr int kCap = kFillCap * TXN;       // sorted candidates per round in LDS
  constexpr int kStage = kStageCap * TXN;    // staged candidates
  constexpr int kNbr = kMaxNeighbours * TXN;  // tiles in reach
  constexpr int kCmap = kChunkMap * TXN;     // chunks mapped to their tile
  // One round sorted by bucket: position words s_sort[0 ..) and key words
  // s_sort[koff ..).  The tiles in reach and their chunk offsets live in the
  // same array above 2 kCap; once the candidates are staged they are dead and
  // a round holds kCapX = kCap + 3 kNbr / 2 candidates (key words from kCapX).
  constexpr int kCapX = kCap + 3 * kNbr / 2;
  static_assert(kNbr % 64 == 0 && kBuckets % 64 == 0, "one-wavefront scans cover whole lanes");
  __shared__ uint32_t s_sort[2 * kCap + 3 * kNbr];
  __shared__ uint32_t s_cnt[kBuckets];          // histogram, then cursors
  __shared__ uint32_t s_start[kBuckets + 1];    // bucket starts over all rounds
  // a tile in reach: its index, and its offset (ix, iy, iz) in the reach box
  // (6 bits each) with its useful list prefix length - 1 above bit 18
  uint32_t* const s_nb_tile = s_sort + 2 * kCap;
  uint32_t* const s_nb_word = s_nb_tile + kNbr;
  // chunk -> tile in reach; while the sub-bricks scan (and when the candidates
  // are staged) it holds each wavefront's compacted candidates instead
  // (chunk -> its tile in reach: tile index | chunk of the list << 24 -- a
  // volume of fewer than 2^32 voxels has fewer than 2^24 tiles -- and the
  // tile's reach word, for the first kDesc chunks)
  constexpr int kDesc = kCmap / 4;
  __shared__ __attribute__((aligned(16))) uint2 s_cmap[kDesc];
  uint32_t* const s_off = s_nb_word + kNbr;  // chunk offsets
  // candidates in list order (kStage position words, then kStage r^2 words;
  // or 2 kStage one-word candidates); when they are not staged, the
  // wavefronts' compacted candidates
  __shared__ __attribute__((aligned(16))) uint32_t s_stage[2 * kStage];
  __shared__ uint32_t s_nn, s_over, s_pending, s_nst, s_nchunks;
  __shared__ unsigned long long s_tgt[8 * TXN];
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  const uint32_t W = I.W, H = I.H, D = I.D, TX = I.TX, TY = I.TY, TZ = I.TZ;
  BM_GLOBAL const uint32_t* __restrict__ lists = gbl(I.lists);
  BM_GLOBAL const TileMeta* __restrict__ meta = gbl(I.meta);
  BM_GLOBAL const unsigned long long* __restrict__ masks = gbl(I.masks);
  BM_GLOBAL PhaseStatus* __restrict__ st = gbl(I.status);
  BM_GLOBAL uint32_t* __restrict__ M = gbl(I.h);  // the Y-pass buffer is free once the Z pass has run
  BM_GLOBAL uint32_t* __restrict__ overflow = gbl(I.overflow);
  constexpr bool diag = kDiag;
  BM_GLOBAL unsigned long long* __restrict__ diag_buf = gbl(I.diag);
  if (!st->has_site) return;
  const unsigned long long clk0 = diag ? wall_clock64() : 0ull;
  unsigned long long clk_p3 = 0, clk_p4 = 0;
  uint32_t t, tx, ty, tz;  // first tile of the block
  if (TXN == 1) {
    const bool grouped = I.tile_order & 2u;
    const uint32_t npos = tile_positions(TX, TY, TZ, grouped);
    const uint32_t jpos = xcd_tile(vb, npos);
    if (jpos >= npos) return;
    if (!walk_of(I, grouped).at(jpos, tx, ty, tz)) return;
    t = (tz * TY + ty) * TX + tx;
  } else {
    const uint32_t TG = (TX + TXN - 1) / TXN;  // groups of TXN tiles along x
    const uint32_t npos = TG * TY * TZ;
    const uint32_t jpos = xcd_tile(vb, npos);
    if (jpos >= npos) return;
    tx = (jpos % TG) * TXN;
    ty = (jpos / TG) % TY;
    tz = jpos / (TG * TY);
    t = (tz * TY + ty) * TX + tx;
  }
  const uint32_t ntx = min((uint32_t)TXN, TX - tx);  // tiles of the block
  const int ox = (int)tx * 8, oy = (int)ty * 8, oz = (int)tz * 8;
  const int hx = min(ox + 8 * TXN - 1, (int)W - 1), hy = min(oy + 7, (int)H - 1),
            hz = min(oz + 7, (int)D - 1);
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t max_r2 = st->max_d2;
  // Ridges of radius above kGatherMaxRadius are left to the scatter pass
  // (k_fill_scatter_big): the gather takes r^2 <= r2cap only.
  const uint32_t r2cap = min(max_r2, kGatherMaxR2);
  const int R = (int)ceil_sqrt_u32(r2cap);
  if (threadIdx.x < 8 * TXN)
    s_tgt[threadIdx.x] = (threadIdx.x >> 3) < ntx
                             ? masks[(size_t)(t + (threadIdx.x >> 3)) * 16 + 8 + (threadIdx.x & 7)]
                             : 0ull;
  if (threadIdx.x == 0) {
    s_nn = 0;
    s_over = 0;
    s_nst = 0;
  }
  // buckets 0 .. gtop hold every candidate (keys r^2 >> gshift of r^2 <= max_r2)
  const uint32_t gshift = bucket_shift(max_r2);
  const uint32_t gtop = max_r2 >> gshift;
  for (uint32_t i = threadIdx.x; i <= gtop; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
  const unsigned long long ta = diag ? wall_clock64() : 0ull;
  bool any_target = false;
  for (int i = 0; i < 8 * TXN; ++i) any_target |= s_tgt[i] != 0;
  const int sx = (int)(wv % (2 * TXN)) * 4, sy = (int)((wv / (2 * TXN)) & 1) * 4,
            sz = (int)(wv / (4 * TXN)) * 4;
  const int vx = sx + (int)(lane & 3), vy = sy + (int)((lane >> 2) & 3), vz = sz + (int)(lane >> 4);
  const uint32_t li = (uint32_t)((vx & 7) + 8 * vy + 64 * vz);  // index within its tile
  const uint32_t tgt_word = 8 * (uint32_t)(vx >> 3) + (li >> 6);
  const bool valid = ox + vx < (int)W && oy + vy < (int)H && oz + vz < (int)D;
  const size_t gidx = ((size_t)(oz + vz) * H + (oy + vy)) * W + (ox + vx);
  if (!any_target) {
    if (valid) M[gidx] = 0;
    return;
  }
  // P1: tiles in reach and the length of their useful list prefix.
  const int K = (R + 7) / 8;
  const uint32_t side = 2 * K + 1, sidex = 2 * K + TXN;
  // sides <= 8 (the common case): 3-bit fields of i instead of divisions
  const bool small = sidex <= 8;
  const uint32_t nnb = small ? 512u : sidex * side * side;
  for (uint32_t i = threadIdx.x; i < nnb; i += blockDim.x) {
    int ix, iy, iz;
    if (small) {
      ix = (int)(i & 7);
      iy = (int)((i >> 3) & 7);
      iz = (int)(i >> 6);
      if (ix >= (int)sidex || iy >= (int)side || iz >= (int)side) continue;
    } else {
      ix = (int)(i % sidex);
      iy = (int)((i / sidex) % side);
      iz = (int)(i / (sidex * side));
    }
    const int ux = (int)tx + ix - K, uy = (int)ty + iy - K, uz = (int)tz + iz - K;
    if (ux < 0 || uy < 0 || uz < 0 || ux >= (int)TX || uy >= (int)TY || uz >= (int)TZ) continue;
    const uint32_t u = ((uint32_t)uz * TY + (uint32_t)uy) * TX + (uint32_t)ux;
    // count and maxr2 in one 8-byte load (TileMeta is 40 bytes, its first two
    // fields adjacent)
    const v2u32 cm = *(BM_GLOBAL const v2u32*)&meta[u];
    const uint32_t mcount = cm.x, mmax = cm.y;
    if (mcount == 0) continue;
    const int ulx = ux * 8, uly = uy * 8, ulz = uz * 8;
    const int uhx = min(ulx + 7, (int)W - 1), uhy = min(uly + 7, (int)H - 1),
              uhz = min(ulz + 7, (int)D - 1);
    const uint32_t gx = ulx > hx ? (uint32_t)(ulx - hx) : (ox > uhx ? (uint32_t)(ox - uhx) : 0u);
    const uint32_t gy = uly > hy ? (uint32_t)(uly - hy) : (oy > uhy ? (uint32_t)(oy - uhy) : 0u);
    const uint32_t gz = ulz > hz ? (uint32_t)(ulz - hz) : (oz > uhz ? (uint32_t)(oz - uhz) : 0u);
    const uint32_t bd2 = gap2(gx, gy, gz);
    if (mmax < bd2) continue;
    const uint32_t len = reach_len(meta + u, mcount, bd2);
    if (len == 0) continue;
    const uint32_t slot = atomicAdd(&s_nn, 1u);
    if (slot < (uint32_t)kNbr) {
      s_nb_tile[slot] = u;
      s_nb_word[slot] = (uint32_t)ix | ((uint32_t)iy << 6) | ((uint32_t)iz << 12) | ((len - 1) << 18);
    } else {
      s_over = 1;
    }
  }
  __syncthreads();
  const unsigned long long tb = diag ? wall_clock64() : 0ull;
  if (s_over) {  // too many tiles in reach: exact slow path
    if (threadIdx.x < ntx) {
      overflow[1 + atomicAdd((uint32_t*)&overflow[0], 1u)] = t + threadIdx.x;
      atomicAdd((uint32_t*)&st->overflow_tiles, 1u);
    }
    return;
  }
  const uint32_t nn = s_nn;
  // chunk offsets of the tiles in reach (64-entry chunks of their prefixes),
  // scanned by one wavefront, which also writes each tile's chunk descriptors
  if (wv == 0) {
    const uint32_t tot = wave_scan_sized<kNbr / 64>(
        nn, lane, [&](uint32_t j) { return ((s_nb_word[j] >> 18) + 64) >> 6; },
        [&](uint32_t j, uint32_t p) {
          s_off[j] = p;
          const uint32_t u = s_nb_tile[j], w = s_nb_word[j];
          const uint32_t c1 = min(p + (((w >> 18) + 64u) >> 6), (uint32_t)kDesc);
          for (uint32_t c = p; c < c1; ++c) s_cmap[c] = make_uint2(u | ((c - p) << 24), w);
        });
    if (lane == 0) s_nchunks = tot;
  }
  __syncthreads();
  const uint32_t nchunks = s_nchunks;
  const unsigned long long tc = diag ? wall_clock64() : 0ull;
  unsigned long long td = tc;
  const uint32_t kshift = gshift + 16;  // bucket of a key word
  // Chunks q = wv, wv + NW, ... of the flattened (neighbour, 64 entries)
  // range; one chunk is a sorted run of one list.  Read kTrip chunks per trip.
  constexpr int kTrip = 4;
  // When every chunk is in the chunk map (the usual case; uniform over the
  // block) a trip has no loop in it, so the LDS lookups of its chunks and then
  // their list loads are all issued together; otherwise a chunk past the map
  // finds its slot by a binary search over the offsets.
  const bool mapped = nchunks <= (uint32_t)kDesc && !(I.flags & kItemNoChunkMap);
  // A list entry of a tile in reach is a candidate if its ball meets the
  // block box.  Positions relative to the block origin, offset by 512 (the
  // packed-position fields): p = (ix - K) * 8 + 512 + the entry's local
  // coordinate, with the tile's part (ix - K) * 8 + 512 uniform over a chunk;
  // the gap to the box [0, w] is max(512 - p, p - (512 + w), 0).
  const int wbx = hx - ox + 512, wby = hy - oy + 512, wbz = hz - oz + 512;
  auto chunk_trips = [&](auto mapped_tag, auto&& use) {
    constexpr bool kMapped = decltype(mapped_tag)::value;
    for (uint32_t q0 = wv; q0 < nchunks; q0 += kTrip * NW) {
      uint32_t ent[kTrip], nbw[kTrip];
#pragma unroll
      for (int k = 0; k < kTrip; ++k) {
        const uint32_t q = q0 + (uint32_t)NW * k;
        // the chunk's values are uniform over the wave: scalar registers
        uint32_t u, w1, e;
        if (kMapped) {
          // one broadcast read of the chunk's descriptor; a chunk q >=
          // nchunks gets e beyond the list's last useful entry
          const uint2 dsc = s_cmap[min(q, nchunks - 1)];
          const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dsc.x);
          w1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dsc.y);
          u = d0 & 0xFFFFFFu;
          e = (q < nchunks ? (d0 >> 24) * 64u : 512u) + lane;
        } else {  // last slot with s_off <= q (offsets increase strictly)
          uint32_t lo = 0, hi = nn - 1;
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= q) lo = mid;
            else hi = mid - 1;
          }
          const uint32_t cur = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
          const uint32_t off = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_off[cur]);
          u = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_nb_tile[cur]);
          w1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_nb_word[cur]);
          e = (q - off) * 64 + lane;
        }
        const uint32_t last = w1 >> 18;  // useful prefix length - 1
        nbw[k] = w1;
        // Every read here is unconditional (the list index clamped to the
        // list's last useful entry), so no branch separates the trip's
        // chunks.  A chunk q >= nchunks has e > last: its slot is the last
        // chunk's, which ends at q - 1 < q.
        const uint32_t v = (lists + (size_t)u * 512)[min(e, last)];
        ent[k] = e <= last ? v : 0u;
      }
#pragma unroll
      for (int k = 0; k < kTrip; ++k) {
        if (q0 + (uint32_t)NW * k >= nchunks) break;  // uniform over the wave
        const uint32_t w1 = nbw[k], en = ent[k];
        const int bx = ((int)(w1 & 63) - K) * 8 + 512, by = ((int)((w1 >> 6) & 63) - K) * 8 + 512,
                  bz = ((int)((w1 >> 12) & 63) - K) * 8 + 512;
        const int px = bx + (int)(en & 7), py = by + (int)((en >> 3) & 7), pz = bz + (int)((en >> 6) & 7);
        const int ex = max(max(512 - px, px - wbx), 0), ey = max(max(512 - py, py - wby), 0),
                  ez = max(max(512 - pz, pz - wbz), 0);
        const uint32_t r2 = en >> 9;
        // (no short circuit: every lane computes the gap, no branch around it)
        const bool keep = (en != 0) & (gap2((uint32_t)ex, (uint32_t)ey, (uint32_t)ez) <= r2) &
                          (r2 <= r2cap);
        const uint32_t pos = (uint32_t)px | ((uint32_t)py << 10) | ((uint32_t)pz << 20);
        use(keep, pos, r2);
      }
    }
  };
  auto chunk_pass = [&](auto&& use) {
    if (mapped) chunk_trips(std::true_type{}, use);
    else chunk_trips(std::false_type{}, use);
  };
  // Staging format.  With R <= 45 a candidate is one word: r^2 (below 2048)
  // above bit 21 and its block-relative centre c + R in three 7-bit fields
  // (c lies in [-R, 8 * TXN - 1 + R], so c + R <= 16 + 90 - 1 < 128); the
  // staging area then holds 2 kStage candidates.  Otherwise a position word
  // and an r^2 word each.
  const bool compact = R <= 45;
  const uint32_t stage_cap = (I.flags & kItemNoStage) ? 0u
                             : compact ? 2u * (uint32_t)kStage : (uint32_t)kStage;
  // P2: histogram by bucket, one LDS atomic per kept candidate (the LDS is
  // otherwise idle here, the vector units are not); the candidates are also
  // staged in LDS (one slot reservation per chunk) so that the placement
  // passes need not read the lists again.
  chunk_pass([&](bool keep, uint32_t pos, uint32_t r2) {
    const unsigned long long km = __builtin_amdgcn_ballot_w64(keep);
    const uint32_t b = keep ? gtop - (r2 >> gshift) : 0u;
    if (keep) atomicAdd(&s_cnt[b], 1u);
    if (km) {
      uint32_t k0 = 0;
      if (lane == 0) k0 = atomicAdd(&s_nst, (uint32_t)__popcll(km));
      k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)k0);
      const uint32_t k = k0 + (uint32_t)__popcll(km & ((1ull << lane) - 1ull));
      if (keep && k < stage_cap) {
        if (compact) {
          s_stage[k] = (r2 << 21) | (((pos >> 20) - 512u + (uint32_t)R) << 14) |
                       ((((pos >> 10) & 1023u) - 512u + (uint32_t)R) << 7) |
                       ((pos & 1023u) - 512u + (uint32_t)R);
        } else {
          s_stage[k] = pos;
          s_stage[kStage + k] = r2;
        }
      }
    }
  });
  __syncthreads();
  const uint32_t nst = s_nst;
  const bool staged = nst <= stage_cap;  // else the placement re-reads the lists
  // candidates per round (a smaller tuned or test capacity is kept as is)
  const uint32_t cap = I.fill_cap >= (uint32_t)kFillCap ? (staged ? (uint32_t)kCapX : (uint32_t)kCap)
                                                        : min(I.fill_cap * TXN, (uint32_t)kCap);
  uint32_t* const s_pos = s_sort;
  uint32_t* const s_r2 = s_sort + (staged ? kCapX : kCap);
  static_assert(sizeof(s_cmap) >= (size_t)NW * 64 * 8 && sizeof(s_stage) >= (size_t)NW * 64 * 8,
                "per-wave candidate buffers");
  // This wave's 64 (position, key) pairs of compacted candidates, in LDS that
  // the scan does not otherwise use: the chunk map once the candidates are
  // staged (the placement then reads the staging area), else the staging area.
  uint2* const wbuf = (staged ? (uint2*)s_cmap : (uint2*)s_stage) + wv * 64;
  // Staged candidates, 64 per wave and trip (the trip count is uniform over a wave).
  auto staged_pass = [&](auto&& use) {
    for (uint32_t i0 = wv * 64; i0 < nst; i0 += blockDim.x) {
      const uint32_t i = i0 + lane;
      const bool keep = i < nst;
      if (compact) {
        const uint32_t w = keep ? s_stage[i] : 0u;
        const uint32_t off = 512u - (uint32_t)R;
        use(keep,
            ((w & 127u) + off) | ((((w >> 7) & 127u) + off) << 10) | ((((w >> 14) & 127u) + off) << 20),
            w >> 21);
      } else {
        use(keep, keep ? s_stage[i] : 0u, keep ? s_stage[kStage + i] : 0u);
      }
    }
  };
  // bucket starts (and the placement cursors), scanned by one wavefront
  if (wv == 0) {
    const uint32_t tot = wave_scan_sized<kBuckets / 64>(
        gtop + 1, lane, [&](uint32_t j) { return s_cnt[j]; },
        [&](uint32_t j, uint32_t p) {
          s_cnt[j] = p;
          s_start[j] = p;
        });
    if (lane == 0) s_start[gtop + 1] = tot;
  }
  __syncthreads();
  const uint32_t n = s_start[gtop + 1];
  const unsigned long long clk1 = diag ? wall_clock64() : 0ull;
  // The sorted candidates are kept as a position word and a key word
  // key = r^2 << 16 | (K & 0xFFFF), and a voxel's running maximum m is the
  // largest key that covers it (r^2 <= 16384 < 2^16 in the gather, so keys
  // order as their r^2 do and m >> 16 is the maximum r^2).  With R <= 64
  // ("dot" form) the position word holds the block-relative centre c as
  // bytes c + 64 (c lies in [-R, 8 * TXN - 1 + R]) and K = |c|^2 - r^2, so
  // that for a voxel v of the block (coordinates below 16)
  //   |v - c|^2 <= r^2  <=>  2 v.c - |v|^2 >= |c|^2 - r^2 = K
  // and 2 v.c - |v|^2 = dot4(2v, c + 64) - (|v|^2 + 128 (vx + vy + vz)):
  // one 4 x u8 dot product per test (|K| < 2^15 for R <= 64).  Otherwise the
  // position word is the 10-bit packed one and K is 0.
  const bool dotp = R <= 64 && !(I.flags & kItemPackedFill);
  const uint32_t v2pk = (uint32_t)(2 * vx) | ((uint32_t)(2 * vy) << 8) | ((uint32_t)(2 * vz) << 16);
  const uint32_t nega =
      0u - (uint32_t)(vx * vx + vy * vy + vz * vz + 128 * (vx + vy + vz));
  auto sorted_words = [&](uint32_t pos, uint32_t r2, uint32_t& pw, uint32_t& kw) {
    if (dotp) {
      const int cx = (int)(pos & 1023) - 512, cy = (int)((pos >> 10) & 1023) - 512,
                cz = (int)(pos >> 20) - 512;
      pw = (uint32_t)(cx + 64) | ((uint32_t)(cy + 64) << 8) | ((uint32_t)(cz + 64) << 16);
      const int k = cx * cx + cy * cy + cz * cz - (int)r2;
      kw = (r2 << 16) | ((uint32_t)k & 0xFFFFu);
    } else {
      pw = pos;
      kw = r2 << 16;
    }
  };
  // Rounds over bucket ranges that fit in LDS, largest radii first.
  const bool target = (s_tgt[tgt_word] >> (li & 63)) & 1ull;
  uint32_t m = 0;
  const unsigned long long tmask = __builtin_amdgcn_ballot_w64(target);
  bool done = tmask == 0;
  uint32_t held = 0xFFFFFFFFu;  // bucket key that must be finished before leaving
  unsigned long long tests = 0, scanned = 0, rounds = 0;
  uint32_t b0 = 0;
  while (b0 <= gtop) {
    ++rounds;
    uint32_t b1 = b0 + 1;
    const uint32_t start0 = s_start[b0];
    if (n - start0 <= cap) {
      b1 = gtop + 1;  // the rest fits: the last round (the usual case)
    } else if (s_start[b1] - start0 > cap) {
      if (threadIdx.x < ntx) {  // one bucket alone exceeds LDS: exact slow path
        overflow[1 + atomicAdd((uint32_t*)&overflow[0], 1u)] = t + threadIdx.x;
        atomicAdd((uint32_t*)&st->overflow_tiles, 1u);
      }
      return;
    } else {  // largest b1 <= gtop + 1 with s_start[b1] - s_start[b0] <= cap (starts are monotone)
      uint32_t lo = b1, hi = gtop + 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (s_start[mid] - start0 <= cap) lo = mid;
        else hi = mid - 1;
      }
      b1 = lo;
    }
    const uint32_t base = start0, cnt_round = s_start[b1] - base;
    const unsigned long long c3 = diag ? wall_clock64() : 0ull;
    // P3: place this round's candidates at their sorted positions (any order
    // inside a bucket: the scan below finishes a bucket before it leaves, and
    // a maximum does not depend on the order of its terms).
    auto place = [&](bool keep0, uint32_t pos, uint32_t r2) {
      const uint32_t b = keep0 ? gtop - (r2 >> gshift) : 0u;
      const bool keep = keep0 && b >= b0 && b < b1;
      if (keep) {
        const uint32_t k = atomicAdd(&s_cnt[b], 1u) - base;
        uint32_t pw, kw;
        sorted_words(pos, r2, pw, kw);
        s_pos[k] = pw;
        s_r2[k] = kw;
      }
    };
    if (staged) staged_pass(place);
    else chunk_pass(place);
    __syncthreads();
    const unsigned long long c4 = diag ? wall_clock64() : 0ull;
    clk_p3 += c4 - c3;
    // P4: this wave's 4x4x4 sub-brick scans the round in order.
    if (!done) {
      unsigned long long wait = (__builtin_amdgcn_uicmp(m, 0u, 32) & tmask);
      for (uint32_t c0 = 0; c0 < cnt_round; c0 += 64) {
        ++scanned;
        const uint32_t ci = c0 + lane;
        // (lanes past the round read its last entry, no branch; the live
        // lane test below keeps them out)
        const bool live = ci < cnt_round;
        const uint32_t cp = s_pos[min(ci, cnt_round - 1)], cr = s_r2[min(ci, cnt_round - 1)];
        // Voxels this chunk can still raise: targets whose maximum so far is
        // below the chunk's largest r^2 (those not covered yet among them).
        // With shift > 0 a bucket mixes radii in any order, so a voxel
        // already covered by a ball of the bucket can still gain a larger one
        // of the same bucket.  The chunk is sorted by bucket, so its largest
        // r^2 is at most the top of its first entry's bucket (one broadcast
        // read instead of a wave reduction).  Box of those voxels (lane =
        // x + 4y + 16z of the sub-brick) from the wave's mask with scalar bit
        // operations; a candidate that misses it cannot change any result.
        // (lane 0 holds entry c0: no LDS round trip)
        const uint32_t first_key = (uint32_t)__builtin_amdgcn_readfirstlane((int)cr) >> kshift;
        const uint32_t cmax = ((first_key + 1u) << gshift) - 1u;
        const unsigned long long open = (__builtin_amdgcn_uicmp(m >> 16, cmax, 36) & tmask);
        bool rel = false;
        if (open) {
          // z from the lowest and highest open lane, y and x from the (y, x)
          // mask of all four z-slices (open != 0: the bit scans are defined)
          const int bz0 = sz + (__builtin_ctzll(open) >> 4), bz1 = sz + ((63 - __builtin_clzll(open)) >> 4);
          const uint32_t t2 = (uint32_t)open | (uint32_t)(open >> 32);
          const uint32_t rows = (t2 | (t2 >> 16)) & 0xFFFFu;  // bit x + 4y
          const int by0 = sy + (__builtin_ctz(rows) >> 2), by1 = sy + ((31 - __builtin_clz(rows)) >> 2);
          const uint32_t r8 = rows | (rows >> 8), xm = (r8 | (r8 >> 4)) & 0xFu;
          const int bx0 = sx + __builtin_ctz(xm), bx1 = sx + 31 - __builtin_clz(xm);
          int cx, cy, cz;
          if (dotp) {
            cx = (int)(cp & 255) - 64;
            cy = (int)((cp >> 8) & 255) - 64;
            cz = (int)((cp >> 16) & 255) - 64;
          } else {
            cx = (int)(cp & 1023) - 512;
            cy = (int)((cp >> 10) & 1023) - 512;
            cz = (int)(cp >> 20) - 512;
          }
          const int ex = max(max(bx0 - cx, cx - bx1), 0), ey = max(max(by0 - cy, cy - by1), 0),
                    ez = max(max(bz0 - cz, cz - bz1), 0);
          rel = live & (gap2((uint32_t)ex, (uint32_t)ey, (uint32_t)ez) <= (cr >> 16)) &
                (held == 0xFFFFFFFFu || (cr >> kshift) == held);
        }
        unsigned long long rm = __builtin_amdgcn_ballot_w64(rel);
        if (dotp) {
          if (rm) {
            // Compact the chunk's relevant candidates into slots 0 .. nrel-1
            // of this wave's LDS buffer (a bijection: the other lanes go
            // after them and carry a key whose K = 0x7FFF no voxel reaches,
            // since 2 v.c - |v|^2 < 2^15), then test them four per trip,
            // each pair read by every lane from the same address (one
            // broadcast 16-byte LDS read per two candidates; the last trip
            // ends at slot 63 at most).  The test is one dot product and a
            // compare.  The wave's LDS operations complete in order, so the
            // reads see this chunk's writes; the previous chunk's reads are
            // complete before the writes, as their values were used.
            const uint32_t nrel = (uint32_t)__popcll(rm);
            const uint32_t below = (uint32_t)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(rm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)rm, 0u));
            const uint32_t dst = rel ? below : nrel + (lane - below);
            wbuf[dst] = make_uint2(rel ? cp : 0u, rel ? cr : 0x7FFFu);
            const uint4* const wq = (const uint4*)wbuf;
            for (uint32_t i = 0; i < nrel; i += 4) {
              const uint4 q0 = wq[i >> 1], q1 = wq[(i >> 1) + 1];
              const uint32_t pk[4] = {q0.x, q0.z, q1.x, q1.z}, key[4] = {q0.y, q0.w, q1.y, q1.w};
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                if ((int)__builtin_amdgcn_udot4(v2pk, pk[u], nega, false) >=
                    (int)(int16_t)(key[u] & 0xFFFFu))
                  m = max(m, key[u]);
              }
            }
            tests += nrel;
          }
        } else {
          while (rm) {
            const int j = __ffsll((unsigned long long)rm) - 1;
            rm &= rm - 1;
            const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)cp, j);
            const uint32_t key = (uint32_t)__builtin_amdgcn_readlane((int)cr, j);
            const int dx = vx - ((int)(p & 1023) - 512), dy = vy - ((int)((p >> 10) & 1023) - 512),
                      dz = vz - ((int)(p >> 20) - 512);
            if ((uint32_t)(dx * dx + dy * dy + dz * dz) <= (key >> 16)) m = max(m, key);
            ++tests;
            if (rm) {  // a second candidate per trip
              const int j2 = __ffsll((unsigned long long)rm) - 1;
              rm &= rm - 1;
              const uint32_t p2 = (uint32_t)__builtin_amdgcn_readlane((int)cp, j2);
              const uint32_t key2 = (uint32_t)__builtin_amdgcn_readlane((int)cr, j2);
              const int ex2 = vx - ((int)(p2 & 1023) - 512),
                        ey2 = vy - ((int)((p2 >> 10) & 1023) - 512),
                        ez2 = vz - ((int)(p2 >> 20) - 512);
              if ((uint32_t)(ex2 * ex2 + ey2 * ey2 + ez2 * ez2) <= (key2 >> 16)) m = max(m, key2);
              ++tests;
            }
          }
        }
        wait = (__builtin_amdgcn_uicmp(m, 0u, 32) & tmask);
        // Leave once every target is covered: candidates come in decreasing
        // r^2, so later ones cannot raise a maximum -- except that a bucket
        // with shift > 0 mixes radii and is finished first (rounds never split
        // a bucket, so it ends in this round).
        if (wait == 0 && gshift == 0) {
          // one r^2 per bucket: a later candidate is never larger
          done = true;
          break;
        }
        if (held == 0xFFFFFFFFu && wait == 0) {
          // the chunk's last entry, held by lane last - c0
          const uint32_t last = min(c0 + 63, cnt_round - 1);
          held = (uint32_t)__builtin_amdgcn_readlane((int)cr, (int)(last - c0)) >> kshift;
        }
        if (held != 0xFFFFFFFFu) {
          const uint32_t last = min(c0 + 63, cnt_round - 1);
          const bool more = last + 1 < cnt_round && (s_r2[last + 1] >> kshift) == held;
          if (!more) {
            done = true;
            break;
          }
        }
      }
    }
    if (diag) clk_p4 += wall_clock64() - c4;
    if (threadIdx.x == 0) s_pending = 0;
    __syncthreads();
    if (!done && lane == 0) s_pending = 1;
    __syncthreads();
    if (!s_pending) break;
    b0 = b1;
    __syncthreads();
  }
  if (valid) M[gidx] = m >> 16;
  // Work counters and phase clocks (diagnostics only): plain stores into this
  // block's slot, summed afterwards by k_reduce_diag.
  if (diag && lane == 0) {
    auto slot = diag_buf + ((size_t)vb * NW + wv) * 16;
    slot[0] = 1;  // waves that reached the end
    slot[1] = tests;
    slot[2] = scanned;
    slot[3] = wv == 0 ? rounds : 0;
    slot[4] = clk1 - clk0;
    slot[5] = clk_p3;
    slot[6] = clk_p4;
    slot[7] = wall_clock64() - clk0;
    slot[8] = ta - clk0;
    slot[9] = tb - ta;
    slot[10] = tc - tb;
    slot[11] = td - tc;
    slot[12] = clk1 - td;
    slot[13] = wv == 0 ? n : 0;
    slot[14] = wv == 0 ? nchunks : 0;
    slot[15] = wv == 0 && !staged ? 1 : 0;
  }
}

__global__ void __launch_bounds__(512, 8)
k_fill_gather(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
              uint32_t nitems) {
  k_fill_gather_t<false, 1>(items, starts, nitems);
}
__global__ void __launch_bounds__(1024, 8)
k_fill_gather2(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
               uint32_t nitems) {
  k_fill_gather_t<false, 2>(items, starts, nitems);
}
__global__ void __launch_bounds__(512)
k_fill_gather_diag(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                   uint32_t nitems) {
  k_fill_gather_t<true, 1>(items, starts, nitems);
}

// Sum of the per-wave diagnostic slots (one block).
__global__ void __launch_bounds__(256) k_reduce_diag(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[blockIdx.x];
  if (!(I.flags & kItemDiag)) return;
  BM_GLOBAL const unsigned long long* __restrict__ diag_buf = gbl(I.diag);
  const uint64_t nslots =
      (uint64_t)((tile_positions(I.TX, I.TY, I.TZ, true) + kXcds - 1) / kXcds * kXcds) * 8;
  BM_GLOBAL PhaseStatus* __restrict__ st = gbl(I.status);
  __shared__ unsigned long long s[16];
  if (threadIdx.x < 16) s[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long acc[16] = {0};
  for (uint64_t i = threadIdx.x; i < nslots; i += blockDim.x)
    for (int k = 0; k < 16; ++k) acc[k] += diag_buf[i * 16 + k];
  for (int k = 0; k < 16; ++k) atomicAdd(&s[k], acc[k]);
  __syncthreads();
  if (threadIdx.x < 16) st->fill_diag16[threadIdx.x] = s[threadIdx.x];
}

// Tiles whose candidate set did not fit in LDS: unsorted exact maximum.
__global__ void __launch_bounds__(512)
k_fill_overflow(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                uint32_t nitems) {
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  const uint32_t W = I.W, H = I.H, D = I.D, TX = I.TX, TY = I.TY, TZ = I.TZ;
  BM_GLOBAL const uint32_t* __restrict__ lists = gbl(I.lists);
  BM_GLOBAL const TileMeta* __restrict__ meta = gbl(I.meta);
  BM_GLOBAL const PhaseStatus* __restrict__ st = gbl(I.status);
  BM_GLOBAL uint32_t* __restrict__ M = gbl(I.h);
  BM_GLOBAL const uint32_t* __restrict__ overflow = gbl(I.overflow);
  if (!st->has_site) return;
  const uint32_t count = overflow[0];
  const uint32_t max_r2 = st->max_d2;
  const int R = (int)ceil_sqrt_u32(min(max_r2, kGatherMaxR2));  // as the gather
  const int K = (R + 7) / 8;
  const uint32_t gshift = bucket_shift(max_r2);
  for (uint32_t oi = vb; oi < count; oi += kOverflowBlocks) {
    const uint32_t t = overflow[1 + oi];
    const uint32_t tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
    const int vx = (int)tx * 8 + (int)(threadIdx.x & 7), vy = (int)ty * 8 + (int)((threadIdx.x >> 3) & 7),
              vz = (int)tz * 8 + (int)(threadIdx.x >> 6);
    if (vx >= (int)W || vy >= (int)H || vz >= (int)D) continue;
    uint32_t m = 0;
    for (int uz = max(0, (int)tz - K); uz <= min((int)TZ - 1, (int)tz + K); ++uz)
      for (int uy = max(0, (int)ty - K); uy <= min((int)TY - 1, (int)ty + K); ++uy)
        for (int ux = max(0, (int)tx - K); ux <= min((int)TX - 1, (int)tx + K); ++ux) {
          const uint32_t u = ((uint32_t)uz * TY + (uint32_t)uy) * TX + (uint32_t)ux;
          const TileMeta mt = *(const TileMeta*)&meta[u];
          if (mt.maxr2 <= m) continue;  // cannot raise the maximum
          for (uint32_t e = 0; e < mt.count; ++e) {
            const uint32_t ent = lists[(size_t)u * 512 + e];
            const uint32_t r2 = ent >> 9;
            if (r2 > kGatherMaxR2) continue;  // scattered by k_fill_scatter_big
            if (r2 <= m) {
              if ((r2 >> gshift) < (m >> gshift)) break;  // sorted: the rest is smaller
              continue;
            }
            const uint32_t l = ent & 511;
            const int dx = vx - (ux * 8 + (int)(l & 7)), dy = vy - (uy * 8 + (int)((l >> 3) & 7)),
                      dz = vz - (uz * 8 + (int)(l >> 6));
            if ((uint32_t)(dx * dx + dy * dy + dz * dz) <= r2) m = r2;
          }
        }
    M[((size_t)vz * H + vy) * W + vx] = m;
  }
}

// Balls of radius above kGatherMaxRadius, after the gather (which leaves them
// out) and its overflow pass: atomic maxima into M.  The ridge pass lists the
// tiles that hold such ridges.  Every block visits every such ridge and takes
// the z-planes -R + vb, -R + vb + kOverflowBlocks, ... of its ball, so one
// large ball is spread over all blocks of the item; a wavefront takes a
// (z, y) row and its lanes consecutive x (coalesced atomics).  Each plane,
// row and run is clipped exactly: |dz| <= isqrt(r^2), |dy| <= isqrt(r^2 -
// dz^2), |dx| <= isqrt(r^2 - dz^2 - dy^2), and to the array.
__global__ void __launch_bounds__(256)
k_fill_scatter_big(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                   uint32_t nitems) {
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  BM_GLOBAL const PhaseStatus* __restrict__ st = gbl(I.status);
  if (!st->has_site || st->max_d2 <= kGatherMaxR2) return;
  const int W = (int)I.W, H = (int)I.H, D = (int)I.D;
  const uint32_t TX = I.TX, TY = I.TY;
  BM_GLOBAL const uint32_t* __restrict__ lists = gbl(I.lists);
  BM_GLOBAL const TileMeta* __restrict__ meta = gbl(I.meta);
  BM_GLOBAL const uint32_t* __restrict__ big = gbl(I.bigtiles);
  BM_GLOBAL uint32_t* __restrict__ M = gbl(I.h);
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6),
            nwv = (int)(blockDim.x >> 6);
  const uint32_t nbig = big[0];
  for (uint32_t bi = 0; bi < nbig; ++bi) {
    const uint32_t t = big[1 + bi];
    const uint32_t count = meta[t].count;
    const int ox = (int)(t % TX) * 8, oy = (int)((t / TX) % TY) * 8, oz = (int)(t / (TX * TY)) * 8;
    for (uint32_t e = 0; e < count; ++e) {
      const uint32_t ent = lists[(size_t)t * 512 + e];
      const uint32_t r2 = ent >> 9, l = ent & 511;
      if (r2 <= kGatherMaxR2) continue;
      const int x = ox + (int)(l & 7), y = oy + (int)((l >> 3) & 7), z = oz + (int)(l >> 6);
      const int R = (int)isqrt_u32(r2);
      for (int dz = -R + (int)vb; dz <= R; dz += (int)kOverflowBlocks) {
        const int nz = z + dz;
        if (nz < 0 || nz >= D) continue;
        const uint32_t rz = r2 - (uint32_t)(dz * dz);
        const int Ry = (int)isqrt_u32(rz);
        for (int dy = -Ry + wv; dy <= Ry; dy += nwv) {
          const int ny = y + dy;
          if (ny < 0 || ny >= H) continue;
          const int Rx = (int)isqrt_u32(rz - (uint32_t)(dy * dy));
          const int x0 = max(x - Rx, 0), x1 = min(x + Rx, W - 1);
          BM_GLOBAL uint32_t* __restrict__ row = M + ((size_t)nz * H + ny) * W;
          for (int nx = x0 + lane; nx <= x1; nx += 64) atomicMax((uint32_t*)&row[nx], r2);
        }
      }
    }
  }
}

// Work of the sphere fill as the reference's propagation does it
// (src/local_thickness.rs:182-201): the sum over ridges of the number of
// voxels of the ball within the array.  One thread per list entry, exact
// integer row extents (diagnostics: runs after the plan).
__global__ void __launch_bounds__(256)
k_fill_work(const ItemDesc* __restrict__ items) {
  const ItemDesc& I = items[0];
  const int W = (int)I.W, H = (int)I.H, D = (int)I.D;
  const uint32_t TX = I.TX, TY = I.TY, ntiles = I.ntiles;
  BM_GLOBAL const uint32_t* __restrict__ lists = gbl(I.lists);
  BM_GLOBAL const TileMeta* __restrict__ meta = gbl(I.meta);
  BM_GLOBAL PhaseStatus* __restrict__ st = gbl(I.status);
  unsigned long long visits = 0;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t count = meta[t].count;
    const int ox = (int)(t % TX) * 8, oy = (int)((t / TX) % TY) * 8, oz = (int)(t / (TX * TY)) * 8;
    for (uint32_t e = threadIdx.x; e < count; e += blockDim.x) {
      const uint32_t ent = lists[(size_t)t * 512 + e];
      const uint32_t r2 = ent >> 9, l = ent & 511;
      const int x = ox + (int)(l & 7), y = oy + (int)((l >> 3) & 7), z = oz + (int)(l >> 6);
      const int R = (int)isqrt_u32(r2);
      for (int nz = max(z - R, 0); nz <= min(z + R, D - 1); ++nz) {
        const uint32_t rz = r2 - (uint32_t)((nz - z) * (nz - z));
        const int Ry = (int)isqrt_u32(rz);
        for (int ny = max(y - Ry, 0); ny <= min(y + Ry, H - 1); ++ny) {
          const int Rx = (int)isqrt_u32(rz - (uint32_t)((ny - y) * (ny - y)));
          visits += (unsigned long long)(min(x + Rx, W - 1) - max(x - Rx, 0) + 1);
        }
      }
    }
  }
  visits = wave_sum_u64(visits);
  if ((threadIdx.x & 63) == 0 && visits) atomicAdd((unsigned long long*)&st->fill_visits, visits);
}

// Total ridge count (diagnostics) from the tile headers.
__global__ void __launch_bounds__(256)
k_count_ridges(const TileMeta* __restrict__ meta, uint32_t ntiles, PhaseStatus* __restrict__ st) {
  unsigned long long c = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x)
    c += meta[t].count;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&st->ridge_count, c);
}

// Ridge flags (diagnostics) from the tile lists.
__global__ void __launch_bounds__(256)
k_lists_to_flags(uint32_t W, uint32_t H, uint32_t D, uint32_t TX, uint32_t TY, uint32_t ntiles,
                 const uint32_t* __restrict__ lists, const TileMeta* __restrict__ meta,
                 uint8_t* __restrict__ flags) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileMeta mt = *(const TileMeta*)&meta[t];
    const int ox = (int)(t % TX) * 8, oy = (int)((t / TX) % TY) * 8, oz = (int)(t / (TX * TY)) * 8;
    for (uint32_t e = threadIdx.x; e < mt.count; e += blockDim.x) {
      const uint32_t l = lists[(size_t)t * 512 + e] & 511;
      const int x = ox + (int)(l & 7), y = oy + (int)((l >> 3) & 7), z = oz + (int)(l >> 6);
      flags[((size_t)z * H + y) * W + x] = 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Clean-up + phase mask + statistics, streamed along z.  A work item is a
// 64 x 8 (x, y) footprint over a chunk of kCleanZ planes; a block walks work
// items (persistent, XCD-contiguous ranges) and, for each, a rolling window
// of planes in LDS:
//   plane s arrives: its 68 x 12 region of M (footprint + 2-voxel halo) and
//     the bit array of its zero voxels;
//   zxy(s): per row of the 66 x 10 one-halo region, the 3 x 3 in-plane
//     dilation of the zeros (bit operations, one thread per 32-bit word);
//     Z3(s-1) = zxy(s-2) | zxy(s-1) | zxy(s): the voxels of plane s-1 with a
//     zero among their 26 neighbours (out-of-array voxels hold kOut, never
//     zero: the array border is not background, src/local_thickness.rs:325);
//   c(s-1): what each one-halo voxel of plane s-1 adds to a neighbour's mean
//     (its diameter if it is positive, inside the array and not on the
//     surface, else 0; src/local_thickness.rs:214-238, 255-268);
//   output plane s-2: the diameter, or the f32 mean of the positive
//     non-surface neighbours in NEIGHBORS_26 order (src/local_thickness.rs:
//     243-277), the phase mask (:31-43), exact sums and the value histogram.
// The rings (M 4 planes, Z3 2, c 3) are sized so that one iteration's output
// phase may overlap the next iteration's loads: three barriers per 512
// outputs, and every M value is read from HBM once per work item (halos
// shared by neighbouring footprints come from L2).
constexpr uint32_t kOut = 0xFFFFFFFFu;  // outside the array: never zero, never summed
constexpr int kRegX = 68, kRegY = 12, kReg = kRegX * kRegY;  // footprint + 2-voxel halo
constexpr int kHalX = 66, kHalY = 10, kHal = kHalX * kHalY;  // footprint + 1-voxel halo
constexpr int kZWords = 30;  // 10 rows x 3 words of one-halo bits

// 1 for a positive contribution, 0 for +0.0f: contributions are +0.0f or
// finite and at least 2 (bits in [2^30, 2^31)), so bit 30 is the answer (a
// shift, where a comparison would become a compare and a carry through VCC).
__device__ __forceinline__ uint32_t nonzero1(float c) { return __float_as_uint(c) >> 30; }

// kLdsDiam: every diameter of the item is in the LDS table (the item's
// largest r^2 is below kDiamLds), so a lookup needs no bounds test.
struct CleanupLds {
  uint32_t m[4][kReg];       // M ring
  unsigned long long zb[14];  // zero bits of the newest plane (816 used, padded)
  uint32_t z3[2][kZWords];   // Z3 ring
  float c[3][kHal];          // contribution ring
  float diam[kDiamLds];      // diameters 2*sqrt(m) for small m
  double rcp[27];            // 1/n rounded to f64, n = 1..26 (rcp[0] unused)
  uint32_t top[kTopBins];    // top-16-bit histogram of the values (quantiles)
};

template <bool kLdsDiam>
__device__ __forceinline__ void
cleanup_stream(const ItemDesc& I, uint32_t vb, CleanupLds& L) {
  auto& s_m = L.m;
  auto& s_zb = L.zb;
  auto& s_z3 = L.z3;
  auto& s_c = L.c;
  auto& s_diam = L.diam;
  auto& s_top = L.top;
  auto& s_rcp = L.rcp;
  BM_GLOBAL const uint32_t* __restrict__ M = gbl(I.h);
  BM_GLOBAL const unsigned long long* __restrict__ masks = gbl(I.masks);
  BM_GLOBAL const float* __restrict__ diam = gbl(I.diam);
  BM_GLOBAL const PhaseStatus* __restrict__ st = gbl(I.status);
  const uint32_t W = I.W, H = I.H, D = I.D, TX = I.TX, TY = I.TY;
  const float scale_f32 = I.scale;
  const int lsb_exp = I.lsb;
  float* __restrict__ map_out = (I.flags & kItemMap) ? I.map : nullptr;
  uint32_t* __restrict__ top16 = (I.flags & kItemQuantiles) ? I.top16 : nullptr;
  if (top16)
    for (uint32_t i = threadIdx.x; i < (uint32_t)kTopBins; i += blockDim.x) s_top[i] = 0;
  if (threadIdx.x < 14) s_zb[threadIdx.x] = 0;  // (the padding word stays 0)
  if (threadIdx.x < 27) s_rcp[threadIdx.x] = threadIdx.x ? 1.0 / (double)threadIdx.x : 0.0;
  unsigned long long cnt = 0;
  u128 s = 0, s2 = 0;
  uint32_t mxb = 0, flags = 0;
  const bool valid_phase = st->has_site != 0;
  const uint32_t nd = valid_phase ? min(st->max_d2 + 1, (uint32_t)kDiamLds) : 0u;
  for (uint32_t i = threadIdx.x; i < nd; i += blockDim.x) s_diam[i] = diam[i];
  auto dia = [&](uint32_t m) {
    return kLdsDiam ? s_diam[m] : m < (uint32_t)kDiamLds ? s_diam[m] : diam[m];
  };
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint32_t nfx = (W + 63) / 64, nfy = (H + 7) / 8;
  const uint32_t npos = valid_phase ? cleanup_work_items(W, H, D) : 0u;
  // Block vb walks positions j = xbase + vb / 8, + G / 8, ... of its XCD's
  // contiguous range (G blocks, a multiple of 8): the blocks of one XCD
  // work on neighbouring footprints at once and share their halos in L2.
  const uint32_t per = (npos + kXcds - 1) / kXcds;
  const uint32_t xbase = (vb % kXcds) * per, xend = min(xbase + per, npos);
  const uint32_t jstep = I.cblocks / kXcds;
  // Region elements of this thread: e0 = t and e1 = t + 512 (< 816); one-halo
  // elements t and t + 512 (< 660); output voxel (lane, wv) of the footprint.
  const bool has1 = t + 512 < (uint32_t)kReg;
  const bool hh1 = t + 512 < (uint32_t)kHal;
  const int ox = (int)lane, oy = (int)wv;
  __syncthreads();
  for (uint32_t j = xbase + vb / kXcds; j < xend; j += jstep) {
    const uint32_t fy = j % nfy, fx = (j / nfy) % nfx, zc = j / (nfy * nfx);
    const int x0 = (int)fx * 64, y0 = (int)fy * 8;
    const int za = (int)(zc * kCleanZ), zb = min(za + (int)kCleanZ, (int)D);
    const int gx = x0 + ox, gy = y0 + oy;
    const bool out_valid = gx < (int)W && gy < (int)H;
    // (the empty asm keeps the element coordinates from being hoisted out of
    // the loops and held in registers)
    auto load = [&](uint32_t e, int z) -> uint32_t {
      asm volatile("" : "+v"(e));
      const int qx = x0 - 2 + (int)(e % kRegX), qy = y0 - 2 + (int)(e / kRegX);
      if (z < 0 || z >= (int)D || qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) return kOut;
      return M[((size_t)z * H + qy) * W + qx];
    };
    // Work items whose region lies inside the array in x and y (the common
    // case): each element is at a fixed offset within a plane, so a load is
    // one add; planes outside [0, D) read kOut.
    const bool inner = x0 >= 2 && y0 >= 2 && x0 + 66 <= (int)W && y0 + 10 <= (int)H;
    const size_t plane = (size_t)W * H;
    uint32_t off0 = 0, off1 = 0;
    if (inner) {
      off0 = (uint32_t)((y0 - 2 + (int)(t / kRegX)) * (int)W + x0 - 2 + (int)(t % kRegX));
      const uint32_t e1 = has1 ? t + 512 : t;
      off1 = (uint32_t)((y0 - 2 + (int)(e1 / kRegX)) * (int)W + x0 - 2 + (int)(e1 % kRegX));
    }
    auto load2 = [&](int z, uint32_t& v0, uint32_t& v1) {
      if (inner) {
        if (z < 0 || z >= (int)D) {
          v0 = v1 = kOut;
        } else {
          BM_GLOBAL const uint32_t* pz = M + (size_t)z * plane;
          v0 = pz[off0];
          v1 = has1 ? pz[off1] : kOut;
        }
      } else {
        v0 = load(t, z);
        v1 = has1 ? load(t + 512, z) : kOut;
      }
    };
    uint32_t pre0, pre1;
    load2(za - 2, pre0, pre1);
    // this voxel's tile mask words: its tile column's (x, y) part, plus the
    // plane's tile layer and index (uniform) per plane
    BM_GLOBAL const unsigned long long* const mcol =
        masks + ((size_t)((uint32_t)gy / 8) * TX + (uint32_t)gx / 8) * 16;
    const size_t layer = (size_t)TY * TX * 16;  // words per tile layer
    uint32_t zxy_a = 0, zxy_b = 0;  // zxy(s-2), zxy(s-1) of this thread's word (t < 30)
    // contribution ring slots of planes sp - 3, sp - 2 and sp - 1 (plane p in
    // slot (p - za + 2) mod 3), rotated once per plane
    uint32_t ca = 0, cb = 1, cc = 2;
    for (int sp = za - 2; sp <= zb + 1; ++sp) {
      const int rel = sp - (za - 2);
      if (rel > 0) {
        const uint32_t t0 = ca;
        ca = cb;
        cb = cc;
        cc = t0;
      }
      // selected voxels of this iteration's output plane sp - 2 (this row of
      // the footprint: 8 tiles), read now so the load is complete by the
      // output phase
      const int pz = sp - 2;
      const unsigned long long selw =
          out_valid && pz >= za ? mcol[(uint32_t)pz / 8 * layer + ((uint32_t)pz & 7)] : 0ull;
      // plane sp: M values and zero bits
      uint32_t* sm = s_m[rel & 3];
      sm[t] = pre0;
      if (has1) sm[t + 512] = pre1;
      {
        const unsigned long long b0 = __ballot(pre0 == 0u);
        const unsigned long long b1 = __ballot(has1 && pre1 == 0u);
        if (lane == 0) {
          s_zb[wv] = b0;
          if (wv < 5) s_zb[8 + wv] = b1;  // words 8..12 (elements 512..831)
        }
      }
      if (sp + 1 <= zb + 1) load2(sp + 1, pre0, pre1);  // the next plane, in flight meanwhile
      __syncthreads();
      if (t < (uint32_t)kZWords) {
        // zxy(sp): bit hx of word k of row hy = a zero among region x in
        // [32k + hx', +2], rows hy .. hy+2 (hx' = bit index)
        const uint32_t hy = t / 3, k = t % 3;
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t dy = 0; dy < 3; ++dy) {
          const uint32_t p = 68u * (hy + dy) + 32u * k;  // first bit
          const unsigned long long lo = s_zb[p >> 6], hi = s_zb[(p >> 6) + 1];
          const uint32_t sh = p & 63;
          const unsigned long long v = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
          const uint32_t v32 = (uint32_t)v, v33 = (uint32_t)(v >> 1), v34 = (uint32_t)(v >> 2);
          acc |= v32 | v33 | v34;
        }
        if (k == 2) acc &= 3u;  // one-halo x < 66
        const uint32_t z3 = zxy_a | zxy_b | acc;  // Z3(sp - 1)
        s_z3[(rel + 1) & 1][t] = z3;              // slot of plane sp - 1
        zxy_a = zxy_b;
        zxy_b = acc;
      }
      __syncthreads();
      if (sp - 1 >= za - 1) {
        // c(sp - 1) over the one-halo region
        const uint32_t* m1 = s_m[(rel - 1) & 3];
        const uint32_t* z3 = s_z3[(rel + 1) & 1];
        float* c1 = s_c[cc];  // slot of plane sp - 1
        auto contrib = [&](uint32_t e) {
          asm volatile("" : "+v"(e));
          const int hx = (int)(e % kHalX), hy = (int)(e / kHalX);
          const uint32_t v = m1[(hy + 1) * kRegX + hx + 1];
          const bool surf = (z3[hy * 3 + (hx >> 5)] >> (hx & 31)) & 1u;
          const bool part = v != 0 && v != kOut && !surf;
          return part ? dia(part ? v : 0u) : 0.0f;
        };
        c1[t] = contrib(t);
        if (hh1) c1[t + 512] = contrib(t + 512);
      }
      __syncthreads();
      if (pz < za) continue;  // no output plane yet
      const bool sel = (selw >> (((uint32_t)gx & 7) + 8 * ((uint32_t)gy & 7))) & 1ull;
      const uint32_t m = s_m[(rel - 2) & 3][(oy + 2) * kRegX + ox + 2];
      float tv = 0.0f;
      if (sel) {
        const uint32_t* z3p = s_z3[rel & 1];  // slot of plane sp - 2
        const int hx = ox + 1, hy = oy + 1;
        const bool surface = m != 0 && ((z3p[hy * 3 + (hx >> 5)] >> (hx & 31)) & 1u);
        if (!surface) {
          tv = dia(m);
        } else {
          // pass 2 (src/local_thickness.rs:243-277): f32 mean of the positive,
          // non-surface neighbours in NEIGHBORS_26 order (two groups of 13
          // reads in flight; non-contributors hold +0.0f, which leaves the f32
          // sum bit for bit as skipping them would)
          const float* cz[3] = {s_c[ca], s_c[cb], s_c[cc]};
          const int c0 = hy * kHalX + hx;
          float sum = 0.0f;
          uint32_t count = 0;
#pragma unroll
          for (int g = 0; g < 26; g += 13) {
            float c[13];
#pragma unroll
            for (int k = 0; k < 13; ++k)
              c[k] = cz[1 + kNb26[g + k][2]][c0 + kNb26[g + k][0] + kHalX * kNb26[g + k][1]];
#pragma unroll
            for (int k = 0; k < 13; ++k) {
              sum += c[k];
              // contributions are +0.0f or at least 2: a positive one has
              // nonzero bits (an integer count, no compare)
              count += nonzero1(c[k]);
            }
            __builtin_amdgcn_sched_barrier(0);
          }
          // sum / count rounded to f32: the f64 product with 1/count rounded to
          // f64 is within 2^-51.9 of the quotient relative, and the quotient
          // lies at least 2^-25/26 from every f32 rounding midpoint (an f32
          // sum over count <= 26 is never one: count times a 25-bit odd
          // significand needs 25 bits; s - count * m is a nonzero multiple of
          // half an ulp of the quotient), so both round alike
          tv = count > 0 ? (float)((double)sum * s_rcp[count]) : dia(m);
        }
        if (tv > 0.0f && tv < INFINITY) acc_value(tv * scale_f32, lsb_exp, cnt, s, s2, mxb, flags);
      }
      if (top16) {
        const uint32_t bin = (__float_as_uint(tv) >> 16) - kTopBase;
        const bool in = tv > 0.0f && bin < (uint32_t)kTopBins;
        if (tv > 0.0f && !in) flags |= 2u;  // outside the histogram's range
        wave_hist_add_runs(s_top, in, in ? bin : 0u);
      }
      if (map_out && out_valid) map_out[((size_t)pz * H + gy) * W + gx] = tv;
    }
    __syncthreads();  // the rings are reused by the next work item
  }
  if (top16) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < (uint32_t)kTopBins; i += blockDim.x)
      if (s_top[i]) atomicAdd(&top16[i], s_top[i]);
  }
  block_reduce_store(cnt, s, s2, mxb, flags, I.partials + vb);
}

__global__ void __launch_bounds__(512, 6)
k_cleanup_stream(const ItemDesc* __restrict__ items, const uint32_t* __restrict__ starts,
                 uint32_t nitems) {
  __shared__ CleanupLds lds;
  uint32_t vb;
  const ItemDesc& I = items[item_of(starts, nitems, blockIdx.x, vb)];
  if (gbl(I.status)->max_d2 < (uint32_t)kDiamLds) cleanup_stream<true>(I, vb, lds);
  else cleanup_stream<false>(I, vb, lds);
}

}  // namespace bm
