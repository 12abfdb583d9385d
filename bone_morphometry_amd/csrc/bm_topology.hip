// bm_topology.hip -- the integer measures of analyze beyond thickness and
// spacing, on the analysis bone A = (bone != 0) inside the ROI
// (bone-morphometry 0.7.0-rc.3, src/topology.rs, src/metrics.rs:134-202).
//
// Everything works on bit rows: a packing pass turns the analysis bone and the
// ROI into one 64-bit word per 64 voxels of a row (bits past the row's end are
// 0), so the measures below handle 64 voxels with a few word operations.
//
// Lattice pass: one thread per (row (y, z), word) of the vertex lattice
// [0..W] x [0..H] x [0..D].  Lattice point (x, y, z) sees the octant of
// voxels (x-1..x, y-1..y, z-1..z); it is a vertex of the complex when any of
// the eight is bone, and the lattice edges leaving it along +x, +y, +z exist
// when any of the four voxels around them is bone (src/topology.rs:53-171).
// The voxel (x, y, z) itself adds its cube, the faces it shares with its -x,
// -y, -z neighbours, its interface faces with them (both inside the ROI,
// one bone; src/topology.rs:442-490) and, when bone, whether a 6-neighbour
// lies outside the ROI or the image (src/metrics.rs:166-202).  The x-1 and
// x+1 columns are the words shifted by one bit, with the neighbouring word's
// end bit carried in; each count is a population count.
//
// BoneJ's edge correction (src/topology.rs:175-427) sums small predicates
// over the six faces, twelve edges and eight corners of the image box; the
// eighteen loops are a table of segments, one thread per loop iteration.
//
// Betti numbers (src/topology.rs:29-51,504-701): beta0 counts 26-connected
// components of A, beta2 the 6-connected components of its complement that
// do not reach the image boundary.  Both come from a union-find whose nodes
// are runs along x, named by their first voxel: a run's start is found from
// the bit row (the bit after the last 0 before it).  A word's runs are united
// with the runs of earlier rows they touch, one union per contact (the bits
// where a contact begins), background runs on the image boundary with an
// extra node N standing for the exterior, and the components are the roots.
// Roots are linked under the smaller index, so the count does not depend on
// the order in which unions land.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "bm_device.h"
#include "bm_engine.h"
#include "bm_topology.h"

namespace bm {

enum Counter : uint32_t {
  kCntVertices, kCntEdges, kCntCubes, kCntShared, kCntInterface, kCntBone, kCntBoundary,
  kCntStackVertices, kCntStackEdges, kCntStackFaces, kCntFaceVertices, kCntFaceEdges,
  kCntEdgeVertices, kCntBeta0, kCntBeta2, kCntExterior, kCounters
};

struct VolumeView {
  const uint8_t* bone;
  const uint8_t* roi;  // null: every voxel inside
  uint32_t W, H, D;
};

// A, the analysis bone, with 0 outside the image (get_signed).
__device__ __forceinline__ uint32_t abone(const VolumeView& v, int64_t x, int64_t y, int64_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= v.W || y >= v.H || z >= v.D) return 0;
  const uint64_t i = ((uint64_t)z * v.H + (uint64_t)y) * v.W + (uint64_t)x;
  return gbl(v.bone)[i] != 0 && (!v.roi || gbl(v.roi)[i] != 0);
}

__device__ __forceinline__ void add_counter(unsigned long long* c, uint32_t k, uint64_t v) {
  if (v) atomicAdd(c + k, (unsigned long long)v);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Adds kN per-thread counts of a 256-thread block into counters[first..]:
// one atomic per counter and block (all blocks hit the same few addresses, so
// per-wave atomics would serialise).
template <int kN>
__device__ __forceinline__ void block_add_counters(const uint64_t (&v)[kN],
                                                   unsigned long long* counters, uint32_t first) {
  __shared__ uint64_t s[4][kN];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kN; ++k) {
    const uint64_t w = wave_sum64(v[k]);
    if (lane == 0) s[wave][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < (uint32_t)kN)
    add_counter(counters, first + threadIdx.x,
                s[0][threadIdx.x] + s[1][threadIdx.x] + s[2][threadIdx.x] + s[3][threadIdx.x]);
}

__global__ void __launch_bounds__(64) k_topo_zero(unsigned long long* c) {
  if (threadIdx.x < kCounters) c[threadIdx.x] = 0;
}

// Bit rows of the analysis bone (A) and of the ROI (R); words per row WW.
struct BitRows {
  const unsigned long long* A;
  const unsigned long long* R;
  uint32_t W, H, D, WW;
};

// Word k of row (y, z), 0 outside the image.
__device__ __forceinline__ unsigned long long row_word(const unsigned long long* bits,
                                                       const BitRows& b, int64_t y, int64_t z,
                                                       int64_t k) {
  if (y < 0 || z < 0 || y >= b.H || z >= b.D || k < 0 || k >= b.WW) return 0;
  return gbl(bits)[((uint64_t)z * b.H + (uint64_t)y) * b.WW + (uint64_t)k];
}

// Packing: one wavefront per row; a ballot per 64 voxels.
__global__ void __launch_bounds__(256)
k_topo_pack(VolumeView v, uint32_t WW, unsigned long long* __restrict__ A,
            unsigned long long* __restrict__ R) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rows = (uint64_t)v.H * v.D;
  const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t row = wave0; row < rows; row += waves) {
    const uint64_t base = row * v.W;
    for (uint32_t k = 0; k < WW; ++k) {
      const uint32_t x = k * 64 + lane;
      uint32_t a = 0, r = 0;
      if (x < v.W) {
        r = !v.roi || gbl(v.roi)[base + x] != 0;
        a = r && gbl(v.bone)[base + x] != 0;
      }
      const unsigned long long wa = __ballot(a), wr = __ballot(r);
      if (lane == 0) {
        gbl(A)[row * WW + k] = wa;
        gbl(R)[row * WW + k] = wr;
      }
    }
  }
}

// The lattice pass (see the file comment): one thread per lattice row and word.
__global__ void __launch_bounds__(256)
k_topo_lattice(BitRows b, unsigned long long* __restrict__ counters) {
  const uint32_t WL = b.W / 64 + 1;  // words covering lattice points 0..W
  const uint64_t total = (uint64_t)(b.H + 1) * (b.D + 1) * WL;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t vert = 0, edges = 0, cubes = 0, shared = 0, iface = 0, bone = 0, boundary = 0;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t k = (int64_t)(g % WL);
    const uint64_t row = g / WL;
    const int64_t y = (int64_t)(row % (b.H + 1)), z = (int64_t)(row / (b.H + 1));
    unsigned long long A[4], Ap[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // q bit 0: +y, bit 1: +z
      const int64_t yy = y - 1 + (q & 1), zz = z - 1 + (q >> 1);
      A[q] = row_word(b.A, b, yy, zz, k);
      Ap[q] = (A[q] << 1) | (row_word(b.A, b, yy, zz, k - 1) >> 63);
    }
    vert += __popcll(A[0] | A[1] | A[2] | A[3] | Ap[0] | Ap[1] | Ap[2] | Ap[3]);
    edges += __popcll(A[0] | A[1] | A[2] | A[3]) + __popcll(Ap[1] | A[1] | Ap[3] | A[3]) +
             __popcll(Ap[2] | A[2] | Ap[3] | A[3]);
    const unsigned long long c = A[3];  // the voxels (x, y, z)
    cubes += __popcll(c);
    shared += __popcll(c & Ap[3]) + __popcll(c & A[2]) + __popcll(c & A[1]);
    const unsigned long long r0 = row_word(b.R, b, y, z, k);
    if (r0 | c) {
      const unsigned long long rm = (r0 << 1) | (row_word(b.R, b, y, z, k - 1) >> 63);
      const unsigned long long rp = (r0 >> 1) | (row_word(b.R, b, y, z, k + 1) << 63);
      const unsigned long long ry = row_word(b.R, b, y - 1, z, k), rz = row_word(b.R, b, y, z - 1, k);
      iface += __popcll(r0 & rm & (c ^ Ap[3])) + __popcll(r0 & ry & (c ^ A[2])) +
               __popcll(r0 & rz & (c ^ A[1]));
      bone += __popcll(c);
      const unsigned long long inside =
          rm & rp & ry & rz & row_word(b.R, b, y + 1, z, k) & row_word(b.R, b, y, z + 1, k);
      boundary += __popcll(c & ~inside);
    }
  }
  // kCntVertices .. kCntBoundary are consecutive
  const uint64_t v[7] = {vert, edges, cubes, shared, iface, bone, boundary};
  block_add_counters<7>(v, counters, kCntVertices);
}

// ---------------------------------------------------------------------------
// BoneJ's stack counts.  A segment is one loop nest of src/topology.rs:203-427:
// each axis is either fixed at {0, n-1} (two values) or runs over [lo, hi).
struct StackSegment {
  uint32_t count;      // iterations
  uint8_t counter;     // kCntStackVertices .. kCntEdgeVertices
  uint8_t formula;     // see stack_term
  uint8_t fixed;       // bit a: axis a (0 x, 1 y, 2 z) is fixed at {0, n-1}
  uint8_t pad;
  int32_t lo[3], len[3];
};
constexpr int kStackSegments = 18;
struct StackTable {
  StackSegment seg[kStackSegments];
};

__device__ __forceinline__ uint32_t stack_term(const VolumeView& v, uint32_t f, int64_t x,
                                               int64_t y, int64_t z) {
  auto F = [&](int64_t a, int64_t b, int64_t c) { return abone(v, a, b, c); };
  switch (f) {
    case 0: return F(x, y, z);
    case 1: return F(x, y, z) | F(x, y - 1, z) | F(x - 1, y - 1, z) | F(x - 1, y, z);
    case 2: return F(x, y, z) | F(x, y - 1, z) | F(x, y - 1, z - 1) | F(x, y, z - 1);
    case 3: return F(x, y, z) | F(x, y, z - 1) | F(x - 1, y, z - 1) | F(x - 1, y, z);
    case 4: return F(x, y, z) ? 2u : F(x, y - 1, z) + F(x - 1, y, z);
    case 5: return F(x, y, z) | F(x, y, z - 1);
    case 6: return F(x, y, z) | F(x - 1, y, z);
    default: return F(x, y, z) | F(x, y - 1, z);
  }
}

__global__ void __launch_bounds__(256)
k_topo_stack(VolumeView v, StackTable t, uint64_t total, unsigned long long* __restrict__ counters) {
  __shared__ unsigned long long s[6];
  if (threadIdx.x < 6) s[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    uint64_t l = g;
    int k = 0;
    while (l >= t.seg[k].count) l -= t.seg[k++].count;
    const StackSegment& sg = t.seg[k];
    const uint32_t n[3] = {v.W, v.H, v.D};
    int64_t c[3];
    for (int a = 0; a < 3; ++a) {
      if (sg.fixed & (1u << a)) {
        c[a] = (l & 1) ? (int64_t)n[a] - 1 : 0;
        l >>= 1;
      } else {
        c[a] = sg.lo[a] + (int64_t)(l % (uint64_t)sg.len[a]);
        l /= (uint64_t)sg.len[a];
      }
    }
    const uint32_t term = stack_term(v, sg.formula, c[0], c[1], c[2]);
    if (term) atomicAdd(&s[sg.counter - kCntStackVertices], (unsigned long long)term);
  }
  __syncthreads();
  if (threadIdx.x < 6) add_counter(counters, kCntStackVertices + threadIdx.x, s[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Union-find over voxel indices (labels[N] is the exterior node).  Loads and
// stores are coherent at device scope; a root is linked under the smaller root.
__device__ __forceinline__ uint32_t uf_load(uint32_t* L, uint32_t i) {
  return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t* L, uint32_t x) {
  uint32_t p = uf_load(L, x);
  while (p != x) {
    const uint32_t g = uf_load(L, p);
    // path splitting: x's parent becomes its grandparent (still an ancestor)
    if (g != p) __hip_atomic_store(L + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = g;
  }
  return x;
}

__device__ void uf_unite(uint32_t* L, uint32_t a, uint32_t b) {
  a = uf_find(L, a);
  b = uf_find(L, b);
  while (a != b) {
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(L + b, b, a);  // b still a root: link it under a
    if (old == b) return;
    b = uf_find(L, old);
    a = uf_find(L, a);
  }
}

// Word k of row (y, z) of the phase being labelled: A, or its complement
// inside the image.  0 outside the image.
__device__ __forceinline__ unsigned long long phase_word(const BitRows& b, bool fg, int64_t y,
                                                         int64_t z, int64_t k) {
  if (y < 0 || z < 0 || y >= b.H || z >= b.D || k < 0 || k >= b.WW) return 0;
  const unsigned long long a = gbl(b.A)[((uint64_t)z * b.H + (uint64_t)y) * b.WW + (uint64_t)k];
  if (fg) return a;
  const uint32_t rest = b.W - (uint32_t)k * 64;  // valid bits in this word
  const unsigned long long valid = rest >= 64 ? ~0ull : ((1ull << rest) - 1);
  return ~a & valid;
}

// First x of the run containing x in row (y, z) (x must be in the phase).
__device__ __forceinline__ uint32_t run_start(const BitRows& b, bool fg, int64_t y, int64_t z,
                                              uint32_t x) {
  int64_t k = x >> 6;
  const uint32_t bit = x & 63;
  unsigned long long gaps = ~phase_word(b, fg, y, z, k) & (~0ull >> (63 - bit));
  while (!gaps) {
    if (--k < 0) return 0;
    gaps = ~phase_word(b, fg, y, z, k);
  }
  return (uint32_t)(k * 64 + 64 - __clzll((long long)gaps));
}

// Run starts in word k of a row: phase bits whose left neighbour is not.
__device__ __forceinline__ unsigned long long starts_of(const BitRows& b, bool fg, int64_t y,
                                                        int64_t z, int64_t k,
                                                        unsigned long long p) {
  return p & ~((p << 1) | (phase_word(b, fg, y, z, k - 1) >> 63));
}

// Labels of the run starts (each its own node) and of the exterior node.
__global__ void __launch_bounds__(256)
k_ccl_init(BitRows b, bool fg, uint32_t* __restrict__ labels) {
  const uint64_t rows = (uint64_t)b.H * b.D, total = rows * b.WW, n = rows * b.W;
  if (blockIdx.x == 0 && threadIdx.x == 0) labels[n] = (uint32_t)n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t k = (int64_t)(g % b.WW);
    const uint64_t row = g / b.WW;
    const int64_t y = (int64_t)(row % b.H), z = (int64_t)(row / b.H);
    unsigned long long s = starts_of(b, fg, y, z, k, phase_word(b, fg, y, z, k));
    while (s) {
      const uint32_t x = (uint32_t)(k * 64 + __ffsll((long long)s) - 1);
      s &= s - 1;
      labels[row * b.W + x] = (uint32_t)(row * b.W + x);
    }
  }
}

// Unions where runs of word k of row (y, z) first touch runs of earlier rows
// (and, for the background, the exterior).  26-connectivity looks at rows
// (y-1, z) and (y-1..y+1, z-1) at x-1..x+1; 6-connectivity at rows (y-1, z),
// (y, z-1) at x.  A run start unites with the neighbour row's runs at x-1, at
// x (when x-1 did not already cover it) and at x+1; a later voxel of the run
// only with a run beginning at x+1 (26) or x (6).
template <bool kConn26>
__global__ void __launch_bounds__(256)
k_ccl_unite(BitRows b, bool fg, uint32_t* __restrict__ labels) {
  const uint64_t rows = (uint64_t)b.H * b.D, total = rows * b.WW, n = rows * b.W;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t k = (int64_t)(g % b.WW);
    const uint64_t row = g / b.WW;
    const int64_t y = (int64_t)(row % b.H), z = (int64_t)(row / b.H);
    const unsigned long long p = phase_word(b, fg, y, z, k);
    if (!p) continue;
    const unsigned long long s = starts_of(b, fg, y, z, k, p);
    auto node = [&](int64_t yy, int64_t zz, uint32_t x) {
      return (uint32_t)(((uint64_t)zz * b.H + (uint64_t)yy) * b.W + run_start(b, fg, yy, zz, x));
    };
    auto unite_bits = [&](unsigned long long u, int64_t yy, int64_t zz, int dx) {
      while (u) {
        const uint32_t x = (uint32_t)(k * 64 + __ffsll((long long)u) - 1);
        u &= u - 1;
        uf_unite(labels, node(y, z, x), node(yy, zz, (uint32_t)((int64_t)x + dx)));
      }
    };
    constexpr int kRows = kConn26 ? 4 : 2;
    const int64_t ry[4] = {y - 1, y - 1, y, y + 1}, rz[4] = {z, z - 1, z - 1, z - 1};
    const int64_t sy[2] = {y - 1, y}, sz[2] = {z, z - 1};
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t yy = kConn26 ? ry[r] : sy[r], zz = kConn26 ? rz[r] : sz[r];
      const unsigned long long q = phase_word(b, fg, yy, zz, k);
      const unsigned long long qlo = (q << 1) | (phase_word(b, fg, yy, zz, k - 1) >> 63);
      if (kConn26) {
        const unsigned long long qhi = (q >> 1) | (phase_word(b, fg, yy, zz, k + 1) << 63);
        unite_bits(s & qlo, yy, zz, -1);
        unite_bits(s & q & ~qlo, yy, zz, 0);
        unite_bits(p & qhi & ~q, yy, zz, 1);
      } else {
        unite_bits(p & q & (s | ~qlo), yy, zz, 0);
      }
    }
    if (!kConn26 && !fg) {  // background runs on the image boundary reach the exterior
      unsigned long long e;
      if (y == 0 || y == (int64_t)b.H - 1 || z == 0 || z == (int64_t)b.D - 1) {
        e = s;
      } else {
        const uint32_t last = b.W - 1;
        e = (k == 0 ? (p & 1ull) : 0ull) |
            ((uint32_t)k == last / 64 ? (p & (1ull << (last & 63))) : 0ull);
      }
      while (e) {
        const uint32_t x = (uint32_t)(k * 64 + __ffsll((long long)e) - 1);
        e &= e - 1;
        uf_unite(labels, node(y, z, x), (uint32_t)n);
      }
    }
  }
}

// Components = run starts that are roots; for the background also whether
// the exterior node was reached.
__global__ void __launch_bounds__(256)
k_ccl_count(BitRows b, bool fg, uint32_t* __restrict__ labels, uint32_t counter,
            unsigned long long* __restrict__ counters) {
  const uint64_t rows = (uint64_t)b.H * b.D, total = rows * b.WW, n = rows * b.W;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t roots = 0;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t k = (int64_t)(g % b.WW);
    const uint64_t row = g / b.WW;
    const int64_t y = (int64_t)(row % b.H), z = (int64_t)(row / b.H);
    unsigned long long s = starts_of(b, fg, y, z, k, phase_word(b, fg, y, z, k));
    while (s) {
      const uint64_t i = row * b.W + (uint64_t)(k * 64 + __ffsll((long long)s) - 1);
      s &= s - 1;
      roots += labels[i] == (uint32_t)i;
    }
  }
  const uint64_t v[1] = {roots};
  block_add_counters<1>(v, counters, counter);
  if (!fg && blockIdx.x == 0 && threadIdx.x == 0 && labels[n] != (uint32_t)n)
    add_counter(counters, kCntExterior, 1);
}

// ---------------------------------------------------------------------------
static StackTable stack_table(const Dims& dm, uint64_t* total) {
  const int64_t w = dm.w, h = dm.h, d = dm.d;
  struct Axis {
    bool fixed;
    int64_t lo, hi;
  };
  auto F = [](void) { return Axis{true, 0, 0}; };
  auto R = [](int64_t lo, int64_t hi) { return Axis{false, lo, hi}; };
  struct Spec {
    Counter counter;
    int formula;
    Axis x, y, z;
  };
  // src/topology.rs:203-427, in source order
  const Spec specs[kStackSegments] = {
      {kCntStackVertices, 0, F(), F(), F()},
      {kCntStackEdges, 0, R(1, w - 1), F(), F()},
      {kCntStackEdges, 0, F(), R(1, h - 1), F()},
      {kCntStackEdges, 0, F(), F(), R(1, d - 1)},
      {kCntStackFaces, 0, R(1, w - 1), R(1, h - 1), F()},
      {kCntStackFaces, 0, R(1, w - 1), F(), R(1, d - 1)},
      {kCntStackFaces, 0, F(), R(1, h - 1), R(1, d - 1)},
      {kCntFaceVertices, 1, R(0, w + 1), R(0, h + 1), F()},
      {kCntFaceVertices, 2, F(), R(0, h + 1), R(1, d)},
      {kCntFaceVertices, 3, R(1, w), F(), R(1, d)},
      {kCntFaceEdges, 4, R(0, w + 1), R(0, h + 1), F()},
      {kCntFaceEdges, 5, R(0, w), F(), R(1, d)},
      {kCntFaceEdges, 6, R(0, w + 1), F(), R(0, d)},
      {kCntFaceEdges, 5, F(), R(0, h), R(1, d)},
      {kCntFaceEdges, 7, F(), R(1, h), R(0, d)},
      {kCntEdgeVertices, 6, R(1, w), F(), F()},
      {kCntEdgeVertices, 7, F(), R(1, h), F()},
      {kCntEdgeVertices, 5, F(), F(), R(1, d)},
  };
  StackTable t;
  std::memset(&t, 0, sizeof t);
  *total = 0;
  for (int k = 0; k < kStackSegments; ++k) {
    const Spec& sp = specs[k];
    StackSegment& sg = t.seg[k];
    sg.counter = (uint8_t)sp.counter;
    sg.formula = (uint8_t)sp.formula;
    const Axis ax[3] = {sp.x, sp.y, sp.z};
    uint64_t count = 1;
    for (int a = 0; a < 3; ++a) {
      if (ax[a].fixed) {
        sg.fixed |= (uint8_t)(1u << a);
        sg.lo[a] = 0;
        sg.len[a] = 2;
        count *= 2;
      } else {
        const int64_t len = std::max<int64_t>(0, ax[a].hi - ax[a].lo);
        sg.lo[a] = (int32_t)ax[a].lo;
        sg.len[a] = (int32_t)std::max<int64_t>(len, 1);
        count *= (uint64_t)len;
      }
    }
    sg.count = (uint32_t)count;
    *total += count;
  }
  return t;
}

void topology_device(Engine& e, const uint8_t* d_bone, const uint8_t* d_roi, const Dims& dims,
                     unsigned parts, TopologyCounts* out, hipStream_t s) {
  const uint64_t n = dims.n();
  if (n >= 0xFFFFFFFFull) throw std::range_error("topology needs fewer than 2^32 - 1 voxels");
  VolumeView v{d_bone, d_roi, dims.w, dims.h, dims.d};
  const uint32_t WW = (dims.w + 63) / 64;
  const uint64_t rows = (uint64_t)dims.h * dims.d;
  const uint64_t words = rows * WW;
  // slot 2: labels (n + 1), then the A and R bit rows
  const size_t label_bytes = ((n + 1) * sizeof(uint32_t) + 255) / 256 * 256;
  char* base = static_cast<char*>(e.scratch_bytes(label_bytes + 2 * words * 8, 2));
  uint32_t* labels = reinterpret_cast<uint32_t*>(base);
  unsigned long long* bitsA = reinterpret_cast<unsigned long long*>(base + label_bytes);
  unsigned long long* bitsR = bitsA + words;
  unsigned long long* d_cnt =
      static_cast<unsigned long long*>(e.scratch_bytes(kCounters * sizeof(unsigned long long), 3));
  BM_LAUNCH("topo_zero", k_topo_zero, 1, 64, 0, s, d_cnt);
  const uint32_t pblocks = (uint32_t)std::min<uint64_t>((rows + 3) / 4, 16384);
  BM_LAUNCH("topo_pack", k_topo_pack, pblocks, 256, 0, s, v, WW, bitsA, bitsR);
  const BitRows b{bitsA, bitsR, dims.w, dims.h, dims.d, WW};
  if (parts & kTopoCubical) {
    const uint64_t total = (uint64_t)(dims.h + 1) * (dims.d + 1) * (dims.w / 64 + 1);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 2048);
    BM_LAUNCH("topo_lattice", k_topo_lattice, blocks, 256, 0, s, b, d_cnt);
  }
  if (parts & kTopoBonej) {
    uint64_t total = 0;
    const StackTable t = stack_table(dims, &total);
    const uint32_t blocks =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 2048));
    BM_LAUNCH("topo_stack", k_topo_stack, blocks, 256, 0, s, v, t, total, d_cnt);
  }
  if (parts & kTopoBetti) {
    const uint32_t wblocks = (uint32_t)std::min<uint64_t>((words + 255) / 256, 16384);
    BM_LAUNCH("ccl_init", k_ccl_init, wblocks, 256, 0, s, b, true, labels);
    const uint32_t cblocks = std::min<uint32_t>(wblocks, 2048);
    BM_LAUNCH("ccl_unite26", k_ccl_unite<true>, wblocks, 256, 0, s, b, true, labels);
    BM_LAUNCH("ccl_count", k_ccl_count, cblocks, 256, 0, s, b, true, labels, (uint32_t)kCntBeta0,
              d_cnt);
    BM_LAUNCH("ccl_init", k_ccl_init, wblocks, 256, 0, s, b, false, labels);
    BM_LAUNCH("ccl_unite6", k_ccl_unite<false>, wblocks, 256, 0, s, b, false, labels);
    BM_LAUNCH("ccl_count", k_ccl_count, cblocks, 256, 0, s, b, false, labels, (uint32_t)kCntBeta2,
              d_cnt);
  }
  BM_HIP(hipGetLastError());
  unsigned long long c[kCounters];
  BM_HIP(hipMemcpyAsync(c, d_cnt, sizeof c, hipMemcpyDeviceToHost, s));
  BM_HIP(hipStreamSynchronize(s));
  out->vertices = (int64_t)c[kCntVertices];
  out->edges = (int64_t)c[kCntEdges];
  out->cubes = (int64_t)c[kCntCubes];
  out->faces = 6 * (int64_t)c[kCntCubes] - (int64_t)c[kCntShared];
  out->interface_faces = (int64_t)c[kCntInterface];
  out->bone_voxels = c[kCntBone];
  out->boundary_bone_voxels = c[kCntBoundary];
  out->stack_vertices = (int64_t)c[kCntStackVertices];
  out->stack_edges = (int64_t)c[kCntStackEdges];
  out->stack_faces = (int64_t)c[kCntStackFaces];
  out->face_vertices = (int64_t)c[kCntFaceVertices];
  out->face_edges = (int64_t)c[kCntFaceEdges];
  out->edge_vertices = (int64_t)c[kCntEdgeVertices];
  out->beta0 = c[kCntBeta0];
  // background components less the one reaching the exterior, if any did
  out->beta2 = c[kCntBeta2] - c[kCntExterior];
}

}  // namespace bm
