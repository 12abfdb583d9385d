// bm_topology.hip -- the integer measures of analyze beyond thickness and
// spacing, on the analysis bone A = (bone != 0) inside the ROI
// (bone-morphometry 0.7.0-rc.3, src/topology.rs, src/metrics.rs:134-202).
//
// Lattice pass: one wavefront per row (y, z) of the vertex lattice
// [0..W] x [0..H] x [0..D].  Lattice point (x, y, z) sees the octant of
// voxels (x-1..x, y-1..y, z-1..z); it is a vertex of the complex when any of
// the eight is bone, and the lattice edges leaving it along +x, +y, +z exist
// when any of the four voxels around them is bone (src/topology.rs:53-171).
// The voxel (x, y, z) itself adds its cube, the faces it shares with its -x,
// -y, -z neighbours, its interface faces with them (both inside the ROI,
// one bone; src/topology.rs:442-490) and, when bone, whether a 6-neighbour
// lies outside the ROI or the image (src/metrics.rs:166-202).  The -x values
// come from the neighbouring lane.
//
// BoneJ's edge correction (src/topology.rs:175-427) sums small predicates
// over the six faces, twelve edges and eight corners of the image box; the
// eighteen loops are a table of segments, one thread per loop iteration.
//
// Betti numbers (src/topology.rs:29-51,504-701): beta0 counts 26-connected
// components of A, beta2 the 6-connected components of its complement that
// do not reach the image boundary.  Both come from a union-find over runs:
// every voxel of a run along x is labelled with the run's first voxel, then
// runs are united where they first touch a run of a previous row (one union
// per contact instead of one per voxel pair), background runs on the image
// boundary are united with an extra node N standing for the exterior, and
// the components are the roots.  Roots are linked smaller index first, so
// the count is independent of the order in which unions land.
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

// Inside the ROI (0 outside the image).
__device__ __forceinline__ uint32_t inroi(const VolumeView& v, int64_t x, int64_t y, int64_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= v.W || y >= v.H || z >= v.D) return 0;
  if (!v.roi) return 1;
  return gbl(v.roi)[((uint64_t)z * v.H + (uint64_t)y) * v.W + (uint64_t)x] != 0;
}

__device__ __forceinline__ void add_counter(unsigned long long* c, uint32_t k, uint64_t v) {
  if (v) atomicAdd(c + k, (unsigned long long)v);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(64) k_topo_zero(unsigned long long* c) {
  if (threadIdx.x < kCounters) c[threadIdx.x] = 0;
}

// The lattice pass (see the file comment); one wavefront per lattice row.
__global__ void __launch_bounds__(256)
k_topo_lattice(VolumeView v, unsigned long long* __restrict__ counters) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rows = (uint64_t)(v.H + 1) * (v.D + 1);
  const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t vert = 0, edges = 0, cubes = 0, shared = 0, iface = 0, bone = 0, boundary = 0;
  for (uint64_t row = wave0; row < rows; row += waves) {
    const int64_t y = (int64_t)(row % (v.H + 1)), z = (int64_t)(row / (v.H + 1));
    // octant rows (y-1|y, z-1|z): A and R at x and (from the lane below) x-1
    uint32_t cA[4] = {0, 0, 0, 0}, cR[4] = {0, 0, 0, 0};  // lane 63 of the previous chunk
    for (int64_t xb = 0; xb <= (int64_t)v.W; xb += 64) {
      const int64_t x = xb + lane;
      uint32_t A[4], R[4], Ap[4], Rp[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {  // k bit 0: +y, bit 1: +z
        const int64_t yy = y - 1 + (k & 1), zz = z - 1 + (k >> 1);
        A[k] = abone(v, x, yy, zz);
        R[k] = inroi(v, x, yy, zz);
        const uint32_t a_up = __shfl_up(A[k], 1), r_up = __shfl_up(R[k], 1);
        Ap[k] = lane ? a_up : cA[k];
        Rp[k] = lane ? r_up : cR[k];
        cA[k] = __shfl(A[k], 63);
        cR[k] = __shfl(R[k], 63);
      }
      // octant bits: a{dx}{dy}{dz} with d = 1 for x, y, z and 0 for x-1, y-1, z-1
      const uint32_t a000 = Ap[0], a100 = A[0], a010 = Ap[1], a110 = A[1];
      const uint32_t a001 = Ap[2], a101 = A[2], a011 = Ap[3], a111 = A[3];
      vert += a000 | a100 | a010 | a110 | a001 | a101 | a011 | a111;
      edges += (a100 | a110 | a101 | a111) + (a010 | a110 | a011 | a111) +
               (a001 | a101 | a011 | a111);
      cubes += a111;
      shared += (a111 & a011) + (a111 & a101) + (a111 & a110);
      // voxel c = (x, y, z): R[3] / A[3]; its -x, -y, -z neighbours
      if (R[3]) iface += (Rp[3] & (a111 ^ a011)) + (R[2] & (a111 ^ a101)) + (R[1] & (a111 ^ a110));
      if (a111) {
        bone += 1;
        const uint32_t inside = Rp[3] & R[2] & R[1] & inroi(v, x + 1, y, z) &
                                inroi(v, x, y + 1, z) & inroi(v, x, y, z + 1);
        boundary += inside ? 0 : 1;
      }
    }
  }
  vert = wave_sum64(vert);
  edges = wave_sum64(edges);
  cubes = wave_sum64(cubes);
  shared = wave_sum64(shared);
  iface = wave_sum64(iface);
  bone = wave_sum64(bone);
  boundary = wave_sum64(boundary);
  if (lane == 0) {
    add_counter(counters, kCntVertices, vert);
    add_counter(counters, kCntEdges, edges);
    add_counter(counters, kCntCubes, cubes);
    add_counter(counters, kCntShared, shared);
    add_counter(counters, kCntInterface, iface);
    add_counter(counters, kCntBone, bone);
    add_counter(counters, kCntBoundary, boundary);
  }
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

// Voxel of the phase being labelled (A == fg).
__device__ __forceinline__ bool in_phase(const VolumeView& v, bool fg, int64_t x, int64_t y,
                                         int64_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= v.W || y >= v.H || z >= v.D) return false;
  return (abone(v, x, y, z) != 0) == fg;
}

// Run labels: each voxel of the phase gets the index of the first voxel of its
// run along x.  One wavefront per row; the exterior node is initialised too.
__global__ void __launch_bounds__(256)
k_ccl_runs(VolumeView v, bool fg, uint32_t* __restrict__ labels) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rows = (uint64_t)v.H * v.D, n = rows * v.W;
  if (blockIdx.x == 0 && threadIdx.x == 0) labels[n] = (uint32_t)n;
  const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t row = wave0; row < rows; row += waves) {
    const int64_t y = (int64_t)(row % v.H), z = (int64_t)(row / v.H);
    const uint64_t base = row * v.W;
    bool open = false;    // a run continues into this chunk
    uint32_t carried = 0;  // its first x
    for (uint32_t xb = 0; xb < v.W; xb += 64) {
      const uint32_t x = xb + lane;
      const bool t = in_phase(v, fg, x, y, z);
      const uint64_t m = __ballot(t);
      const uint64_t starts = m & ~((m << 1) | (open ? 1ull : 0ull));
      const uint64_t le = starts & (~0ull >> (63 - lane));
      const uint32_t first = le ? xb + 63 - (uint32_t)__clzll((long long)le) : carried;
      if (t) labels[base + x] = (uint32_t)(base + first);
      if ((m >> 63) & 1) {
        if (starts) carried = xb + 63 - (uint32_t)__clzll((long long)starts);
        open = true;
      } else {
        open = false;
      }
    }
  }
}

// Unions where runs first touch a run of an earlier row (and, for the
// background, the exterior).  26-connectivity looks at rows (y-1, z) and
// (y-1..y+1, z-1) at x-1..x+1; 6-connectivity at rows (y-1, z), (y, z-1) at x.
template <bool kConn26>
__global__ void __launch_bounds__(256)
k_ccl_unite(VolumeView v, bool fg, uint32_t* __restrict__ labels) {
  const uint64_t n = (uint64_t)v.W * v.H * v.D;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t x = (int64_t)(i % v.W);
    const uint64_t yz = i / v.W;
    const int64_t y = (int64_t)(yz % v.H), z = (int64_t)(yz / v.H);
    if (!in_phase(v, fg, x, y, z)) continue;
    const bool start = !in_phase(v, fg, x - 1, y, z);
    auto idx = [&](int64_t a, int64_t b, int64_t c) {
      return (uint32_t)(((uint64_t)c * v.H + (uint64_t)b) * v.W + (uint64_t)a);
    };
    if (kConn26) {
      const int64_t ry[4] = {y - 1, y - 1, y, y + 1}, rz[4] = {z, z - 1, z - 1, z - 1};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool lo = in_phase(v, fg, x - 1, ry[r], rz[r]);
        const bool mid = in_phase(v, fg, x, ry[r], rz[r]);
        const bool hi = in_phase(v, fg, x + 1, ry[r], rz[r]);
        if (start) {
          if (lo) uf_unite(labels, (uint32_t)i, idx(x - 1, ry[r], rz[r]));
          if (mid && !lo) uf_unite(labels, (uint32_t)i, idx(x, ry[r], rz[r]));
          if (hi && !mid) uf_unite(labels, (uint32_t)i, idx(x + 1, ry[r], rz[r]));
        } else if (hi && !mid) {
          uf_unite(labels, (uint32_t)i, idx(x + 1, ry[r], rz[r]));
        }
      }
    } else {
      const int64_t ry[2] = {y - 1, y}, rz[2] = {z, z - 1};
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (in_phase(v, fg, x, ry[r], rz[r]) && (start || !in_phase(v, fg, x - 1, ry[r], rz[r])))
          uf_unite(labels, (uint32_t)i, idx(x, ry[r], rz[r]));
      const bool edge = x == 0 || x == (int64_t)v.W - 1 || y == 0 || y == (int64_t)v.H - 1 ||
                        z == 0 || z == (int64_t)v.D - 1;
      if (!fg && edge && (start || x == (int64_t)v.W - 1)) uf_unite(labels, (uint32_t)i, (uint32_t)n);
    }
  }
}

// Components = roots among the phase's voxels; for the background also
// whether the exterior node was reached.
__global__ void __launch_bounds__(256)
k_ccl_count(VolumeView v, bool fg, uint32_t* __restrict__ labels, uint32_t counter,
            unsigned long long* __restrict__ counters) {
  const uint64_t n = (uint64_t)v.W * v.H * v.D;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t roots = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t x = (int64_t)(i % v.W);
    const uint64_t yz = i / v.W;
    const int64_t y = (int64_t)(yz % v.H), z = (int64_t)(yz / v.H);
    if (in_phase(v, fg, x, y, z) && labels[i] == (uint32_t)i) ++roots;
  }
  roots = wave_sum64(roots);
  if ((threadIdx.x & 63) == 0) add_counter(counters, counter, roots);
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
  unsigned long long* d_cnt =
      static_cast<unsigned long long*>(e.scratch_bytes(kCounters * sizeof(unsigned long long), 3));
  BM_LAUNCH("topo_zero", k_topo_zero, 1, 64, 0, s, d_cnt);
  if (parts & kTopoCubical) {
    const uint64_t rows = (uint64_t)(dims.h + 1) * (dims.d + 1);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((rows + 3) / 4, 8192);
    BM_LAUNCH("topo_lattice", k_topo_lattice, blocks, 256, 0, s, v, d_cnt);
  }
  if (parts & kTopoBonej) {
    uint64_t total = 0;
    const StackTable t = stack_table(dims, &total);
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 2048));
    BM_LAUNCH("topo_stack", k_topo_stack, blocks, 256, 0, s, v, t, total, d_cnt);
  }
  if (parts & kTopoBetti) {
    uint32_t* labels = static_cast<uint32_t*>(e.scratch_bytes((n + 1) * sizeof(uint32_t), 2));
    const uint64_t rows = (uint64_t)dims.h * dims.d;
    const uint32_t rblocks = (uint32_t)std::min<uint64_t>((rows + 3) / 4, 8192);
    const uint32_t vblocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
    BM_LAUNCH("ccl_runs", k_ccl_runs, rblocks, 256, 0, s, v, true, labels);
    BM_LAUNCH("ccl_unite", k_ccl_unite<true>, vblocks, 256, 0, s, v, true, labels);
    BM_LAUNCH("ccl_count", k_ccl_count, vblocks, 256, 0, s, v, true, labels, (uint32_t)kCntBeta0,
              d_cnt);
    BM_LAUNCH("ccl_runs", k_ccl_runs, rblocks, 256, 0, s, v, false, labels);
    BM_LAUNCH("ccl_unite", k_ccl_unite<false>, vblocks, 256, 0, s, v, false, labels);
    BM_LAUNCH("ccl_count", k_ccl_count, vblocks, 256, 0, s, v, false, labels, (uint32_t)kCntBeta2,
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
