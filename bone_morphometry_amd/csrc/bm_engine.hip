// bm_engine.hip -- host side of the device engine: arenas, launch sequences.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "bm_engine.h"
#include "bm_tiles.h"

namespace bm {
constexpr uint32_t kXcds = 8;  // must match bm_device.h
}

namespace bm {

// Kernels (bm_kernels.hip).
__global__ void k_edt_x(const uint8_t*, const uint8_t*, uint32_t, uint16_t*, uint32_t, uint32_t,
                        BlockPartial*);
__global__ void k_reduce_status(const BlockPartial*, uint32_t, const uint32_t*, uint32_t,
                                PhaseStatus*);
__global__ void k_edt_y(const uint16_t*, uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t);
__global__ void k_edt_z(const uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t,
                        uint32_t, uint32_t*);
__global__ void k_luts(const PhaseStatus*, uint32_t, uint32_t*, float*);
// Tile kernels (bm_tiles.hip).
__global__ void k_ridges_tiled(const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                               uint32_t, const uint32_t*, uint32_t, PhaseStatus*, uint32_t*,
                               TileMeta*, unsigned long long*);
__global__ void k_fill_gather(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                              const uint32_t*, const TileMeta*, const unsigned long long*,
                              PhaseStatus*, uint32_t*, uint32_t*, uint32_t, uint32_t,
                              unsigned long long*);
__global__ void k_reduce_diag(const unsigned long long*, uint64_t, PhaseStatus*);
__global__ void k_count_ridges(const TileMeta*, uint32_t, PhaseStatus*);
__global__ void k_fill_overflow(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                const uint32_t*, const TileMeta*, const PhaseStatus*, uint32_t*,
                                const uint32_t*);
__global__ void k_fill_scatter_lists(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                     const uint32_t*, const TileMeta*, const PhaseStatus*,
                                     uint32_t*);
__global__ void k_zero_if_scatter(uint32_t*, uint64_t, const PhaseStatus*);
__global__ void k_lists_to_flags(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                 const uint32_t*, const TileMeta*, uint8_t*);
__global__ void k_cleanup_tiled(const uint32_t*, const unsigned long long*, const float*,
                                const PhaseStatus*, uint32_t, uint32_t, uint32_t, uint32_t,
                                uint32_t, uint32_t, float, int, float*, StatsPartial*, uint32_t*);
__global__ void k_stats_final(const StatsPartial*, uint32_t, StatsPartial*);
__global__ void k_fill_u32(uint32_t*, uint64_t, uint32_t);
__global__ void k_fill_u8(uint8_t*, uint64_t, uint8_t);
__global__ void k_phase_inside(const uint8_t*, const uint8_t*, uint64_t, uint32_t, uint8_t*);

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw DeviceError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  }
}

Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

void Arena::reserve(size_t bytes) {
  if (bytes <= cap_) return;
  if (base_) {
    BM_HIP(hipDeviceSynchronize());
    BM_HIP(hipFree(base_));
    base_ = nullptr;
    cap_ = 0;
  }
  size_t want = bytes + bytes / 8;  // headroom against small growth
  BM_HIP(hipMalloc(&base_, want));
  cap_ = want;
}

Engine::Engine(int device) : device_(device) {
  BM_HIP(hipSetDevice(device));
  BM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Engine::~Engine() {
  if (stream_) (void)hipStreamDestroy(stream_);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

PhaseBuffers Engine::buffers_for(const Dims& dims) {
  const uint64_t n = dims.n();
  const uint64_t bound = dims.max_d2();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const size_t o_dx = take(n * sizeof(uint16_t));
  const size_t o_h = take(n * sizeof(uint32_t));
  const size_t o_d2 = take(n * sizeof(uint32_t));
  const uint32_t TX = tiles_of(dims.w), TY = tiles_of(dims.h), TZ = tiles_of(dims.d);
  const uint64_t ntiles = (uint64_t)TX * TY * TZ;
  const size_t o_l = take(ntiles * 512 * sizeof(uint32_t));
  const size_t o_meta = take(ntiles * sizeof(TileMeta));
  const size_t o_mask = take(ntiles * 16 * sizeof(unsigned long long));
  const size_t o_of = take((ntiles + 1) * sizeof(uint32_t));
  const size_t o_t = take(3 * (bound + 1) * sizeof(uint32_t));
  const size_t o_dm = take((bound + 1) * sizeof(float));
  const size_t o_st = take(sizeof(PhaseStatus));
  const size_t o_xp = take(kEdtXBlocks * sizeof(BlockPartial));
  // Z-pass blocks for any z chunk of at least kMinZChunk lines
  const uint64_t zblocks = (uint64_t)((dims.w + 31) / 32) * dims.h * ((dims.d + kMinZChunk - 1) / kMinZChunk);
  const size_t o_zm = take(zblocks * sizeof(uint32_t));
  const size_t o_p = take(kStatsBlocks * sizeof(StatsPartial));
  const size_t o_s = take(sizeof(StatsPartial));
  const size_t o_top = take(kTopBins * sizeof(uint32_t));
  work_.reserve(off);
  char* b = static_cast<char*>(work_.base());
  PhaseBuffers pb;
  pb.dx = reinterpret_cast<uint16_t*>(b + o_dx);
  pb.h = reinterpret_cast<uint32_t*>(b + o_h);
  pb.d2 = reinterpret_cast<uint32_t*>(b + o_d2);
  pb.lists = reinterpret_cast<uint32_t*>(b + o_l);
  pb.meta = reinterpret_cast<TileMeta*>(b + o_meta);
  pb.masks = reinterpret_cast<unsigned long long*>(b + o_mask);
  pb.overflow = reinterpret_cast<uint32_t*>(b + o_of);
  pb.tiles[0] = TX;
  pb.tiles[1] = TY;
  pb.tiles[2] = TZ;
  pb.tmpl = reinterpret_cast<uint32_t*>(b + o_t);
  pb.diam = reinterpret_cast<float*>(b + o_dm);
  pb.status = reinterpret_cast<PhaseStatus*>(b + o_st);
  pb.xpart = reinterpret_cast<BlockPartial*>(b + o_xp);
  pb.zmax = reinterpret_cast<uint32_t*>(b + o_zm);
  pb.partials = reinterpret_cast<StatsPartial*>(b + o_p);
  pb.stats = reinterpret_cast<StatsPartial*>(b + o_s);
  pb.top16 = reinterpret_cast<uint32_t*>(b + o_top);
  pb.bound = bound;
  return pb;
}

void* Engine::scratch_bytes(size_t bytes, int slot) {
  stage_[slot].reserve(bytes);
  return stage_[slot].base();
}

static unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

void Engine::run_edt(const PhaseSource& src, const Dims& dims, const PhaseBuffers& b,
                     uint32_t* d2, hipStream_t s, const Tuning& t) {
  const uint32_t W = dims.w, H = dims.h, D = dims.d;
  // status <- 0 (a small fill kernel, never a memset node: graph-safe)
  k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(b.status),
                              sizeof(PhaseStatus) / sizeof(uint32_t), 0u);
  const uint32_t rows = H * D;
  const unsigned xblocks = grid_for((uint64_t)rows * 64, 256, kEdtXBlocks);
  BM_LAUNCH("edt_x", k_edt_x, xblocks, 256, 0, s, src.data, src.roi, src.want_nonzero, b.dx, W,
            rows, b.xpart);
  {
    const uint32_t chunk = H <= t.y_rows ? H : t.y_rows - 2 * t.y_halo;
    const uint32_t halo = H <= t.y_rows ? 0 : t.y_halo;
    const uint32_t rows_lds = std::min<uint32_t>(H, chunk + 2 * halo);
    dim3 grid((W + 63) / 64, D, (H + chunk - 1) / chunk);
    BM_LAUNCH("edt_y", k_edt_y, grid, 256, rows_lds * 64 * sizeof(uint16_t), s, b.dx, b.h, W, H,
              chunk, halo);
  }
  {
    const uint32_t chunk = D <= t.z_rows ? D : t.z_rows - 2 * t.z_halo;
    const uint32_t halo = D <= t.z_rows ? 0 : t.z_halo;
    const uint32_t rows_lds = std::min<uint32_t>(D, chunk + 2 * halo);
    dim3 grid((W + 31) / 32, H, (D + chunk - 1) / chunk);
    if (chunk < kMinZChunk && chunk < D) throw std::invalid_argument("z tile chunk below kMinZChunk");
    BM_LAUNCH("edt_z", k_edt_z, grid, 256, rows_lds * 32 * sizeof(uint32_t), s, b.h, d2, W, H, D,
              chunk, halo, b.zmax);
    BM_LAUNCH("reduce_status", k_reduce_status, 1, 1024, 0, s, b.xpart, xblocks, b.zmax,
              grid.x * grid.y * grid.z, b.status);
  }
  BM_HIP(hipGetLastError());
}

void Engine::run_phase_inside(const uint8_t* bone, const uint8_t* roi, uint64_t n,
                              bool select_bone, uint8_t* out, hipStream_t s) {
  k_phase_inside<<<grid_for(n, 256 * 8, 65536), 256, 0, s>>>(bone, roi, n, select_bone ? 1u : 0u,
                                                             out);
  BM_HIP(hipGetLastError());
}

void Engine::run_thickness(const Dims& dims, const PhaseBuffers& b, float* map_out,
                           float scale_f32, int lsb_exp, hipStream_t s, uint8_t* ridge_flags_out,
                           uint32_t* propagated_out, const Tuning& tune) {
  const uint32_t W = dims.w, H = dims.h, D = dims.d;
  const uint64_t n = dims.n();
  const uint32_t bound = (uint32_t)b.bound;
  BM_LAUNCH("luts", k_luts, grid_for((uint64_t)bound + 1, 256, 1u << 30), 256, 0, s, b.status,
            bound, b.tmpl, b.diam);
  const uint32_t TX = b.tiles[0], TY = b.tiles[1], TZ = b.tiles[2];
  const uint32_t ntiles = TX * TY * TZ;
  const uint32_t xgrid = ((ntiles + kXcds - 1) / kXcds) * kXcds;  // see xcd_tile
  BM_LAUNCH("ridges", k_ridges_tiled, xgrid, 512, 0, s, b.d2, W, H, D, TX, TY, TZ, b.tmpl, bound,
            b.status, b.lists, b.meta, b.masks);
  uint32_t* M = b.h;  // the Y-pass buffer is free once the Z pass has run
  k_fill_u32<<<1, 64, 0, s>>>(b.overflow, 1, 0u);
  // The gather fills every tile when the largest radius is at most
  // kGatherMaxRadius; otherwise it returns at once and the scatter runs.
  uint32_t cap = tune.fill_cap ? std::min<uint32_t>(tune.fill_cap, kFillCap) : kFillCap;
  if (const char* env = std::getenv("BM_FILL_CAP")) {  // tuning / test override
    const long v = std::strtol(env, nullptr, 10);
    if (v >= 1 && v <= kFillCap) cap = (uint32_t)v;
  }
  const uint32_t diag = std::getenv("BM_FILL_DIAG") ? 1u : 0u;  // work counters
  unsigned long long* diag_buf = nullptr;
  if (diag) {
    diag_buf = static_cast<unsigned long long*>(scratch_bytes((size_t)xgrid * 8 * 16 * 8, 2));
    k_fill_u32<<<grid_for((uint64_t)xgrid * 8 * 32, 256, 65536), 256, 0, s>>>(
        reinterpret_cast<uint32_t*>(diag_buf), (uint64_t)xgrid * 8 * 32, 0u);
  }
  BM_LAUNCH("fill_gather", k_fill_gather, xgrid, 512, 0, s, W, H, D, TX, TY, TZ, b.lists, b.meta,
            b.masks, b.status, M, b.overflow, cap, diag, diag_buf);
  if (diag) k_reduce_diag<<<1, 256, 0, s>>>(diag_buf, (uint64_t)xgrid * 8, b.status);
  BM_LAUNCH("fill_overflow", k_fill_overflow, 256, 512, 0, s, W, H, D, TX, TY, TZ, b.lists,
            b.meta, b.status, M, b.overflow);
  BM_LAUNCH("fill_scatter_zero", k_zero_if_scatter, grid_for(n, 256, 65536), 256, 0, s, M, n,
            b.status);
  BM_LAUNCH("fill_scatter", k_fill_scatter_lists, 2048, 256, 0, s, W, H, D, TX, TY, ntiles, b.lists,
            b.meta, b.status, M);
  if (ridge_flags_out) {
    k_count_ridges<<<256, 256, 0, s>>>(b.meta, ntiles, b.status);
    k_fill_u8<<<grid_for(n, 256, 65536), 256, 0, s>>>(ridge_flags_out, n, 0);
    k_lists_to_flags<<<4096, 256, 0, s>>>(W, H, D, TX, TY, ntiles, b.lists, b.meta,
                                          ridge_flags_out);
  }
  if (propagated_out) BM_HIP(hipMemcpyAsync(propagated_out, M, n * 4, hipMemcpyDeviceToDevice, s));
  uint32_t* top16 = map_out ? b.top16 : nullptr;  // quantiles need the map and its histogram
  if (top16) k_fill_u32<<<8, 256, 0, s>>>(top16, kTopBins, 0u);
  BM_LAUNCH("cleanup_stats", k_cleanup_tiled, kStatsBlocks, 512, 0, s, M, b.masks, b.diam,
            b.status, W, H, D, TX, TY, ntiles, scale_f32, lsb_exp, map_out, b.partials, top16);
  BM_LAUNCH("stats_final", k_stats_final, 1, 256, 0, s, b.partials, kStatsBlocks, b.stats);
  BM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Engines live until the process ends and are never destroyed: freeing device
// memory from a static destructor could run after the HIP runtime is gone.
static std::mutex g_engines_mu;
static std::map<int, Engine*>& engines() {
  static std::map<int, Engine*>* m = new std::map<int, Engine*>();
  return *m;
}

int current_device() {
  int dev = 0;
  BM_HIP(hipGetDevice(&dev));
  return dev;
}

Engine& engine_for_device(int device) {
  std::lock_guard<std::mutex> g(g_engines_mu);
  auto& m = engines();
  auto it = m.find(device);
  if (it != m.end()) return *it->second;
  Engine* e = new Engine(device);
  m.emplace(device, e);
  return *e;
}

// ---------------------------------------------------------------------------
// Kernel profiling.
namespace {
struct ProfRecord {
  std::string name;
  hipEvent_t start, end;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRecord>* g_prof_records = new std::vector<ProfRecord>();
std::vector<hipEvent_t>* g_prof_pool = new std::vector<hipEvent_t>();

hipEvent_t take_event() {
  if (!g_prof_pool->empty()) {
    hipEvent_t e = g_prof_pool->back();
    g_prof_pool->pop_back();
    return e;
  }
  hipEvent_t e;
  BM_HIP(hipEventCreate(&e));
  return e;
}
}  // namespace

void profile_enable(bool on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_on = on;
}

bool profile_enabled() { return g_prof_on; }

hipEvent_t profile_begin() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  return take_event();
}

void profile_end(const char* name, hipEvent_t start, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  hipEvent_t end = take_event();
  BM_HIP(hipEventRecord(end, s));
  g_prof_records->push_back({name, start, end});
}

int profile_read(KernelTime* out, int max) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  std::vector<std::pair<std::string, std::pair<uint64_t, double>>> agg;
  for (auto& r : *g_prof_records) {
    BM_HIP(hipEventSynchronize(r.end));
    float ms = 0.0f;
    BM_HIP(hipEventElapsedTime(&ms, r.start, r.end));
    bool found = false;
    for (auto& a : agg)
      if (a.first == r.name) {
        a.second.first += 1;
        a.second.second += ms;
        found = true;
        break;
      }
    if (!found) agg.push_back({r.name, {1, (double)ms}});
    g_prof_pool->push_back(r.start);
    g_prof_pool->push_back(r.end);
  }
  g_prof_records->clear();
  int n = 0;
  for (auto& a : agg) {
    if (n >= max) break;
    std::memset(out[n].name, 0, sizeof out[n].name);
    std::strncpy(out[n].name, a.first.c_str(), sizeof out[n].name - 1);
    out[n].launches = a.second.first;
    out[n].total_ms = a.second.second;
    ++n;
  }
  return n;
}

// ---------------------------------------------------------------------------
// Exact statistics on the host.
int stats_lsb_exp(float scale_f32) {
  int e = 0;
  (void)std::frexp(scale_f32, &e);  // scale = f * 2^e, f in [0.5, 1)
  return (e - 1) - 26;
}

// Correctly rounded (ties-to-even) conversion of a 128-bit integer to f64.
static double u128_to_double(unsigned __int128 v) {
  if (v == 0) return 0.0;
  int nbits = 128 - (v >> 64 ? __builtin_clzll((unsigned long long)(v >> 64))
                             : 64 + __builtin_clzll((unsigned long long)v));
  if (nbits <= 53) return (double)(unsigned long long)v;
  const int drop = nbits - 53;
  unsigned __int128 top = v >> drop;
  const unsigned __int128 rem = v - (top << drop);
  const unsigned __int128 half = (unsigned __int128)1 << (drop - 1);
  if (rem > half || (rem == half && (top & 1))) top += 1;
  return std::ldexp((double)(unsigned long long)top, drop);  // top <= 2^53: exact
}

StatsResult finish_stats(const StatsPartial& p, int lsb_exp) {
  StatsResult r;
  r.count = p.count;
  const unsigned __int128 s = ((unsigned __int128)p.sum_hi << 64) | p.sum_lo;
  const unsigned __int128 s2 = ((unsigned __int128)p.sq_hi << 64) | p.sq_lo;
  r.sum = std::ldexp(u128_to_double(s), lsb_exp);
  r.sum_of_squares = std::ldexp(u128_to_double(s2), 2 * lsb_exp);
  float mx;
  std::memcpy(&mx, &p.max_bits, 4);
  r.maximum = (double)mx;
  r.flags = p.flags;
  return r;
}

}  // namespace bm
