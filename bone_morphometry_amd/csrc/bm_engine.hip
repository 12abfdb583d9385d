// bm_engine.hip -- host side of the device engine: plans (item layouts, start
// tables, launch sequences), engines, kernel profiling, exact statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "bm_device.h"
#include "bm_engine.h"
#include "bm_tiles.h"

namespace bm {

// Kernels (bm_kernels.hip).
__global__ void k_prepare(const ItemDesc*);
__global__ void k_edt_x(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_edt_y(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_edt_z(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_reduce_status(const ItemDesc*);
__global__ void k_luts(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_stats_final(const ItemDesc*);
__global__ void k_fill_u32(uint32_t*, uint64_t, uint32_t);
__global__ void k_fill_u8(uint8_t*, uint64_t, uint8_t);
__global__ void k_phase_inside(const uint8_t*, const uint8_t*, uint64_t, uint32_t, uint8_t*);
__global__ void k_collect(const ItemDesc*, ItemResult*);
// Tile kernels (bm_tiles.hip).
__global__ void k_ridges_stream(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_fill_gather(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_fill_gather_diag(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_fill_gather2(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_reduce_diag(const ItemDesc*);
__global__ void k_fill_overflow(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_fill_scatter_big(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_fill_work(const ItemDesc*);
__global__ void k_cleanup_stream(const ItemDesc*, const uint32_t*, uint32_t);
__global__ void k_count_ridges(const TileMeta*, uint32_t, PhaseStatus*);
__global__ void k_lists_to_flags(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                 const uint32_t*, const TileMeta*, uint8_t*);
// Quantile kernels (bm_quantiles.hip).
__global__ void k_select_init(const ItemDesc*);
__global__ void k_top16_pick(const ItemDesc*);
__global__ void k_digit_hist(const ItemDesc*, const uint32_t*, uint32_t, int);
__global__ void k_digit_pick(const ItemDesc*, int);

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

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static uint32_t grid_for(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

static uint32_t round8(uint64_t v) { return (uint32_t)((v + kXcds - 1) / kXcds * kXcds); }

// Rows an EDT Y/Z tile stages: [lo & ~1, hi) rounded up to whole pairs, for a
// line of n rows cut in chunks (even when split) with halos of the given size.
static uint32_t pair_rows(uint32_t n, uint32_t chunk, uint32_t halo) {
  const uint32_t span = std::min(n, chunk + 2 * halo);  // hi - lo
  // an odd halo puts lo on an odd row whenever it is not clipped at 0
  return ((halo & 1u) ? span + 2 : span + 1) & ~1u;
}

// Largest ridge r^2 the tile lists hold, plus one (kListR2Limit).  Tests lower
// it with BM_LIST_R2_LIMIT to reach the out-of-range error on small volumes.
static uint32_t list_r2_limit() {
  uint32_t lim = kListR2Limit;
  if (const char* env = std::getenv("BM_LIST_R2_LIMIT")) {
    const long x = std::strtol(env, nullptr, 10);
    if (x >= 1 && x <= (long)kListR2Limit) lim = (uint32_t)x;
  }
  return lim;
}

// ---------------------------------------------------------------------------
// Item layout: geometry, block counts and buffer offsets (relative to the
// item's base when base is null).  Returns the bytes the item needs.
static size_t lay_out(const ItemSpec& s, ItemDesc& d, char* base, uint32_t* blocks,
                      uint32_t fill_tiles = 1) {
  const Dims& dims = s.dims;
  std::memset(&d, 0, sizeof d);
  d.data = s.src.data;
  d.roi = s.src.roi;
  d.want_nonzero = s.src.want_nonzero;
  d.W = dims.w;
  d.H = dims.h;
  d.D = dims.d;
  d.n = dims.n();
  d.TX = tiles_of(dims.w);
  d.TY = tiles_of(dims.h);
  d.TZ = tiles_of(dims.d);
  d.ntiles = d.TX * d.TY * d.TZ;
  d.bound = (uint32_t)dims.max_d2();
  d.scale = s.scale;
  d.lsb = s.lsb;
  const bool lt = s.thickness;
  const bool quant = lt && s.quantiles;
  const bool mapped = lt && (s.want_map || quant || s.map_out);
  d.flags = (lt ? kItemThickness : 0u) | (quant ? kItemQuantiles : 0u) |
            (lt && s.diag ? kItemDiag : 0u) | (mapped ? kItemMap : 0u);
  const Tuning& t = s.tune;
  if (lt && !t.fill_dot) d.flags |= kItemPackedFill;
  if (lt && !t.fill_chunk_map) d.flags |= kItemNoChunkMap;
  if (lt && !t.fill_stage) d.flags |= kItemNoStage;
  d.xblocks = grid_for((uint64_t)dims.h * dims.d * 64, 256, kEdtXBlocks);
  {
    const uint32_t H = dims.h;
    // split columns take even chunks: the Y pass stages and searches pairs of
    // rows from an even row
    d.ychunk = H <= t.y_rows ? H : t.y_rows > 2 * t.y_halo + 1 ? (t.y_rows - 2 * t.y_halo) & ~1u : 0u;
    d.yhalo = H <= t.y_rows ? 0 : t.y_halo;
    if (d.ychunk == 0) throw std::invalid_argument("y tile chunk below 2 rows");
    d.ygx = (dims.w + 63) / 64;
    d.ygy = dims.d;
    d.ygz = (H + d.ychunk - 1) / d.ychunk;
  }
  {
    const uint32_t D = dims.d;
    d.zchunk = D <= t.z_rows ? D : t.z_rows > 2 * t.z_halo ? (t.z_rows - 2 * t.z_halo) & ~1u : 0u;
    d.zhalo = D <= t.z_rows ? 0 : t.z_halo;
    if (d.zchunk < kMinZChunk && d.zchunk < D)
      throw std::invalid_argument("z tile chunk below kMinZChunk");
    d.zgx = (dims.w + 31) / 32;
    d.zgy = dims.h;
    d.zgz = (D + d.zchunk - 1) / d.zchunk;
  }
  d.lblocks = lt ? std::min<uint32_t>(std::max<uint32_t>((d.bound + 4) / 4, 1), 128) : 0;
  d.fill_cap = t.fill_cap ? std::min<uint32_t>(t.fill_cap, kFillCap) : kFillCap;
  d.tile_order = t.tile_order & 7u;
  d.r2_limit = list_r2_limit();
  {
    const uint32_t dv[4] = {d.TX, d.TY, (d.TX + 3) / 4, (d.TY + 3) / 4};
    for (int i = 0; i < 4; ++i) {
      const FastDiv f(dv[i]);
      d.divs[i][0] = f.d;
      d.divs[i][1] = f.m;
      d.divs[i][2] = f.s1;
      d.divs[i][3] = f.s2;
    }
  }
  const uint32_t cmax = t.cleanup_blocks
                            ? std::min<uint32_t>(std::max<uint32_t>(t.cleanup_blocks & ~7u, 8),
                                                 kStatsBlocks)
                            : (uint32_t)kStatsBlocks;
  d.cblocks =
      lt ? std::min<uint32_t>(std::max<uint32_t>(round8(cleanup_work_items(d.W, d.H, d.D)), 8), cmax)
         : 0;
  d.qblocks = quant ? std::min<uint32_t>(std::max<uint32_t>((uint32_t)(d.n / 32768), 8), 1024) : 0;
  // contiguous ranges of whole 4096-value trips
  d.qper = quant ? ((d.n + d.qblocks - 1) / d.qblocks + 4095) / 4096 * 4096 : 0;
  if (blocks) {
    blocks[kStEdtX] = d.xblocks;
    blocks[kStEdtY] = d.ygx * d.ygy * d.ygz;
    blocks[kStEdtZ] = d.zgx * d.zgy * d.zgz;
    blocks[kStLuts] = d.lblocks;
    // one block per position of the ridge and fill orders (the larger of the two)
    blocks[kStRidges] = lt ? round8(ridge_work_items(d.W, d.H, d.D)) : 0;
    blocks[kStTiles] = !lt ? 0
                       : fill_tiles == 2
                           ? round8((uint64_t)((d.TX + 1) / 2) * d.TY * d.TZ)
                           : round8(tile_positions(d.TX, d.TY, d.TZ, t.tile_order & 2u));
    blocks[kStOverflow] = lt ? kOverflowBlocks : 0;
    blocks[kStCleanup] = d.cblocks;
    blocks[kStQHist] = d.qblocks;
  }
  // Buffers.
  const uint64_t n = d.n;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return base + o;
  };
  d.status = reinterpret_cast<PhaseStatus*>(take(sizeof(PhaseStatus)));
  d.xpart = reinterpret_cast<BlockPartial*>(take((size_t)d.xblocks * sizeof(BlockPartial)));
  d.zmax = reinterpret_cast<uint32_t*>(take((size_t)d.zgx * d.zgy * d.zgz * sizeof(uint32_t)));
  d.dx = reinterpret_cast<uint16_t*>(take(n * sizeof(uint16_t)));
  d.h = reinterpret_cast<uint32_t*>(take(n * sizeof(uint32_t)));
  d.d2 = s.d2_out ? s.d2_out : reinterpret_cast<uint32_t*>(take(n * sizeof(uint32_t)));
  if (lt) {
    d.lists = reinterpret_cast<uint32_t*>(take((size_t)d.ntiles * 512 * sizeof(uint32_t)));
    d.meta = reinterpret_cast<TileMeta*>(take((size_t)d.ntiles * sizeof(TileMeta)));
    d.masks = reinterpret_cast<unsigned long long*>(take((size_t)d.ntiles * 16 * 8));
    d.overflow = reinterpret_cast<uint32_t*>(take(((size_t)d.ntiles + 1) * sizeof(uint32_t)));
    d.bigtiles = reinterpret_cast<uint32_t*>(take(((size_t)d.ntiles + 1) * sizeof(uint32_t)));
    d.tmpl = reinterpret_cast<uint32_t*>(take(3 * ((size_t)d.bound + 1) * sizeof(uint32_t)));
    d.diam = reinterpret_cast<float*>(take(((size_t)d.bound + 1) * sizeof(float)));
    d.partials = reinterpret_cast<StatsPartial*>(take((size_t)d.cblocks * sizeof(StatsPartial)));
    d.stats = reinterpret_cast<StatsPartial*>(take(sizeof(StatsPartial)));
    if (mapped) {
      if (s.map_out) d.map = s.map_out;
      else if (!s.keep_d2 && !s.d2_out) d.map = reinterpret_cast<float*>(d.d2);  // d2 is dead by then
      else d.map = reinterpret_cast<float*>(take(n * sizeof(float)));
    }
    if (quant) {
      d.top16 = reinterpret_cast<uint32_t*>(take(kTopBins * sizeof(uint32_t)));
      d.sel = reinterpret_cast<SelectState*>(take(sizeof(SelectState)));
      d.qhist = reinterpret_cast<unsigned long long*>(take(kMaxRanks * 256 * 8));
      d.qcnt = reinterpret_cast<uint32_t*>(take((size_t)d.qblocks * sizeof(uint32_t)));
      // the propagated map is dead once the clean-up has run (items with
      // quantiles have no stage read-back)
      d.cand = d.h;
    }
    if (d.flags & kItemDiag)
      d.diag = reinterpret_cast<unsigned long long*>(
          take((size_t)round8(tile_positions(d.TX, d.TY, d.TZ, true)) * 8 * 16 * 8));
  }
  return off;
}

size_t Plan::bytes_for(const std::vector<ItemSpec>& specs) {
  size_t total = 0;
  for (const auto& s : specs) {
    ItemDesc d;
    total += lay_out(s, d, nullptr, nullptr);
  }
  return total;
}

size_t Plan::meta_bytes_for(size_t nit) {
  const size_t o_starts = align_up(nit * sizeof(ItemDesc), 256);
  const size_t o_res = align_up(o_starts + (size_t)kNumStages * (nit + 1) * sizeof(uint32_t), 256);
  return align_up(o_res + nit * sizeof(ItemResult), 256);
}

Plan::~Plan() {
  if (h_results_) (void)hipHostFree(h_results_);
}

void Plan::build(const std::vector<ItemSpec>& specs, hipStream_t s, void* ext, size_t ext_cap,
                 bool ext_meta) {
  const size_t nit = specs.size();
  if (nit == 0) throw std::invalid_argument("a plan needs at least one item");
  desc_.assign(nit, ItemDesc{});
  std::vector<size_t> offs(nit);
  std::vector<uint32_t> blocks(nit * kNumStages);
  size_t total = 0;
  any_ = 0;
  y_lds_ = z_lds_ = 0;
  fill_tiles_ = specs[0].tune.fill_tiles == 2 ? 2u : 1u;
  for (const auto& s : specs)
    if (s.thickness && s.diag) fill_tiles_ = 1;  // the diagnostic fill has one tile per block
  for (size_t i = 0; i < nit; ++i) {
    offs[i] = total;
    total += lay_out(specs[i], desc_[i], nullptr, &blocks[i * kNumStages], fill_tiles_);
    any_ |= desc_[i].flags;
    // 16-bit values in pairs of rows from an even row
    y_lds_ = std::max(y_lds_, pair_rows(desc_[i].H, desc_[i].ychunk, desc_[i].yhalo) * 64u * 2u);
    z_lds_ = std::max(z_lds_, pair_rows(desc_[i].D, desc_[i].zchunk, desc_[i].zhalo) * 32u * 2u);
  }
  if (y_lds_ > 65536 || z_lds_ > 65536) throw std::invalid_argument("EDT tile exceeds 64 KiB of LDS");
  char* base;
  ext_meta_ = ext && ext_meta;
  const size_t meta_bytes = ext_meta_ ? meta_bytes_for(nit) : 0;
  if (ext) {
    if (total + meta_bytes > ext_cap || total + meta_bytes < total)
      throw std::range_error("plan exceeds the memory given to it");
    base = static_cast<char*>(ext);
  } else {
    mem_.reserve(total);
    base = static_cast<char*>(mem_.base());
  }
  bytes_ = total;
  for (size_t i = 0; i < nit; ++i)
    lay_out(specs[i], desc_[i], base + offs[i], nullptr, fill_tiles_);
  // Start tables: stage-major, nit + 1 entries each.
  starts_.assign((size_t)kNumStages * (nit + 1), 0);
  for (int st = 0; st < kNumStages; ++st) {
    uint64_t acc = 0;
    for (size_t i = 0; i < nit; ++i) {
      starts_[st * (nit + 1) + i] = (uint32_t)acc;
      acc += blocks[i * kNumStages + st];
    }
    if (acc > 0x7FFFFFFFull) throw std::range_error("too many blocks in one launch");
    starts_[st * (nit + 1) + nit] = (uint32_t)acc;
    total_[st] = (uint32_t)acc;
  }
  const size_t o_desc = 0;
  const size_t o_starts = align_up(nit * sizeof(ItemDesc), 256);
  const size_t o_res = align_up(o_starts + starts_.size() * sizeof(uint32_t), 256);
  char* mb;
  if (ext_meta_) {
    mb = base + align_up(total, 256);
  } else {
    reserve_items(nit);
    mb = static_cast<char*>(meta_.base());
  }
  d_desc_ = reinterpret_cast<ItemDesc*>(mb + o_desc);
  d_starts_ = reinterpret_cast<uint32_t*>(mb + o_starts);
  d_results_ = reinterpret_cast<ItemResult*>(mb + o_res);
  BM_HIP(hipMemcpyAsync(d_desc_, desc_.data(), nit * sizeof(ItemDesc), hipMemcpyHostToDevice, s));
  BM_HIP(hipMemcpyAsync(d_starts_, starts_.data(), starts_.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice, s));
  BM_HIP(hipStreamSynchronize(s));
  results_.assign(nit, ItemResult{});
}

void Plan::reserve_items(size_t nit) {
  const size_t o_starts = align_up(nit * sizeof(ItemDesc), 256);
  const size_t o_res = align_up(o_starts + (size_t)kNumStages * (nit + 1) * sizeof(uint32_t), 256);
  meta_.reserve(o_res + nit * sizeof(ItemResult));
  if (h_results_cap_ < nit) {
    if (h_results_) BM_HIP(hipHostFree(h_results_));
    h_results_ = nullptr;
    BM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_results_), nit * sizeof(ItemResult)));
    h_results_cap_ = nit;
  }
}

void Plan::launch(hipStream_t s) const {
  const uint32_t nit = (uint32_t)desc_.size();
  auto st = [&](int stage) { return d_starts_ + (size_t)stage * (nit + 1); };
  k_prepare<<<nit, 256, 0, s>>>(d_desc_);
  BM_LAUNCH("edt_x", k_edt_x, total_[kStEdtX], 256, 0, s, d_desc_, st(kStEdtX), nit);
  BM_LAUNCH("edt_y", k_edt_y, total_[kStEdtY], 256, y_lds_, s, d_desc_, st(kStEdtY), nit);
  BM_LAUNCH("edt_z", k_edt_z, total_[kStEdtZ], 256, z_lds_, s, d_desc_, st(kStEdtZ), nit);
  BM_LAUNCH("reduce_status", k_reduce_status, nit, 1024, 0, s, d_desc_);
  if (any_ & kItemThickness) {
    BM_LAUNCH("luts", k_luts, total_[kStLuts], 256, 0, s, d_desc_, st(kStLuts), nit);
    BM_LAUNCH("ridges", k_ridges_stream, total_[kStRidges], 512, 0, s, d_desc_, st(kStRidges),
              nit);
    if (any_ & kItemDiag)
      BM_LAUNCH("fill_gather", k_fill_gather_diag, total_[kStTiles], 512, 0, s, d_desc_,
                st(kStTiles), nit);
    else if (fill_tiles_ == 2)
      BM_LAUNCH("fill_gather", k_fill_gather2, total_[kStTiles], 1024, 0, s, d_desc_,
                st(kStTiles), nit);
    else
      BM_LAUNCH("fill_gather", k_fill_gather, total_[kStTiles], 512, 0, s, d_desc_,
                st(kStTiles), nit);
    if (any_ & kItemDiag) k_reduce_diag<<<nit, 256, 0, s>>>(d_desc_);
    BM_LAUNCH("fill_overflow", k_fill_overflow, total_[kStOverflow], 512, 0, s, d_desc_,
              st(kStOverflow), nit);
    BM_LAUNCH("fill_scatter", k_fill_scatter_big, total_[kStOverflow], 256, 0, s, d_desc_,
              st(kStOverflow), nit);
    BM_LAUNCH("cleanup_stats", k_cleanup_stream, total_[kStCleanup], 512, 0, s, d_desc_,
              st(kStCleanup), nit);
    BM_LAUNCH("stats_final", k_stats_final, nit, 256, 0, s, d_desc_);
    if (any_ & kItemQuantiles) {
      BM_LAUNCH("quantile_pick", k_select_init, nit, 256, 0, s, d_desc_);
      BM_LAUNCH("quantile_pick", k_top16_pick, nit, 256, 0, s, d_desc_);
      for (int shift = 8; shift >= 0; shift -= 8) {
        BM_LAUNCH("quantile_hist", k_digit_hist, total_[kStQHist], 256, 0, s, d_desc_,
                  st(kStQHist), nit, shift);
        BM_LAUNCH("quantile_pick", k_digit_pick, nit, 256, 0, s, d_desc_, shift);
      }
    }
  }
  k_collect<<<nit, 64, 0, s>>>(d_desc_, d_results_);
  BM_HIP(hipGetLastError());
}

void Plan::fetch(hipStream_t s) {
  const size_t nit = desc_.size();
  if (ext_meta_) {  // pageable destination: no pinned buffer to allocate
    BM_HIP(hipMemcpyAsync(results_.data(), d_results_, nit * sizeof(ItemResult),
                          hipMemcpyDeviceToHost, s));
    BM_HIP(hipStreamSynchronize(s));
    return;
  }
  BM_HIP(hipMemcpyAsync(h_results_, d_results_, nit * sizeof(ItemResult), hipMemcpyDeviceToHost,
                        s));
  BM_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < nit; ++i) results_[i] = h_results_[i];
}

// ---------------------------------------------------------------------------
Engine::Engine(int device) : device_(device) {
  BM_HIP(hipSetDevice(device));
  BM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Engine::~Engine() {
  if (stream_) (void)hipStreamDestroy(stream_);
}

void* Engine::scratch_bytes(size_t bytes, int slot) {
  stage_[slot].reserve(bytes);
  return stage_[slot].base();
}

void Engine::run_phase_inside(const uint8_t* bone, const uint8_t* roi, uint64_t n,
                              bool select_bone, uint8_t* out, hipStream_t s) {
  k_phase_inside<<<grid_for(n, 256 * 8, 65536), 256, 0, s>>>(bone, roi, n, select_bone ? 1u : 0u,
                                                             out);
  BM_HIP(hipGetLastError());
}

void Engine::stage_outputs(const Plan& p, uint8_t* ridge_flags_out, uint32_t* propagated_out,
                           hipStream_t s) {
  const ItemDesc& d = p.desc(0);
  if (!(d.flags & kItemThickness)) return;
  k_fill_work<<<4096, 256, 0, s>>>(p.device_desc());
  if (ridge_flags_out) {
    k_count_ridges<<<256, 256, 0, s>>>(d.meta, d.ntiles, d.status);
    k_fill_u8<<<grid_for(d.n, 256, 65536), 256, 0, s>>>(ridge_flags_out, d.n, 0);
    k_lists_to_flags<<<4096, 256, 0, s>>>(d.W, d.H, d.D, d.TX, d.TY, d.ntiles, d.lists, d.meta,
                                          ridge_flags_out);
  }
  if (propagated_out)
    BM_HIP(hipMemcpyAsync(propagated_out, d.h, d.n * 4, hipMemcpyDeviceToDevice, s));
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
bool g_prof_suspended = false;
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

void profile_suspend(bool on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_suspended = on;
}

bool profile_enabled() { return g_prof_on && !g_prof_suspended; }

// ",name1,name2," (empty: every kernel)
static std::string& prof_only() {
  static std::string* s = new std::string();
  return *s;
}

void profile_only(const char* names) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  prof_only() = names && *names ? std::string(",") + names + "," : std::string();
}

bool profile_enabled_for(const char* name) {
  if (!profile_enabled()) return false;
  std::lock_guard<std::mutex> g(g_prof_mu);
  const std::string& only = prof_only();
  return only.empty() || only.find(std::string(",") + name + ",") != std::string::npos;
}

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
