// bm_engine.h -- device engine for the exact squared EDT and local thickness.
//
// Everything here is stream-ordered device work on one GPU.  The unit of work
// is an *item*: one phase (bone or marrow, or a Phase of a BinaryVolume) of one
// volume.  A Plan holds the device buffers and descriptors of a list of items
// and launches each stage of the pipeline once for all of them; the kernels
// find their item from the block index.  The one-volume calls of the
// reference's API are plans of one item (EDT, local thickness) or two items
// (the bone and marrow phases of analyze); the batched and serving forms are
// plans of many items.  The host-facing mirror of the reference crate lives in
// bm_api.cpp, the C-ABI in include/bm.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace bm {

// Error carrying the reference's message text (anyhow::bail! strings).
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};
struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

void hip_check(hipError_t e, const char* what);
#define BM_HIP(call) ::bm::hip_check((call), #call)

// Sentinels of the EDT passes.
constexpr uint32_t kInf16 = 0xFFFFu;      // X pass: no site in the row
constexpr uint32_t kInf32 = 0x40000000u;  // Y/Z passes: no site reachable

// Which voxels are selected: roi[i] != 0 (or no ROI) and
// (data[i] != 0) == want_nonzero.  Phase::Foreground is want_nonzero = 1,
// Phase::Background want_nonzero = 0 (src/edt.rs:17-24); the bone and marrow
// phases of analyze are (bone, roi, 1) and (bone, roi, 0) (src/volume.rs:101).
struct PhaseSource {
  const uint8_t* data = nullptr;
  const uint8_t* roi = nullptr;
  uint32_t want_nonzero = 1;
};

struct Dims {
  uint32_t w = 0, h = 0, d = 0;
  uint64_t n() const { return (uint64_t)w * h * d; }
  uint64_t max_d2() const {
    return (uint64_t)(w - 1) * (w - 1) + (uint64_t)(h - 1) * (h - 1) + (uint64_t)(d - 1) * (d - 1);
  }
};

// Per-item device status, written by the kernels of the item.
struct PhaseStatus {
  unsigned long long selected;     // voxels of the phase (X pass)
  unsigned long long ridge_count;  // distance ridges found (stage outputs only)
  unsigned long long fill_visits;  // sphere-fill work (stage outputs only)
  uint32_t has_site;               // any non-selected voxel (X pass)
  uint32_t max_d2;                 // maximum squared distance (Z pass)
  uint32_t error;                  // bit 0: r^2 at or above the item's r2_limit
  uint32_t overflow_tiles;         // tiles whose fill candidates exceeded LDS
  unsigned long long fill_diag16[16];  // fill work counters and clocks (diagnostics)
};

// Per-block partials of the X pass.
struct BlockPartial {
  unsigned long long selected;
  uint32_t flag;
  uint32_t pad;
};

// Exact statistics accumulator: fixed point with LSB 2^lsb_exp.
struct StatsPartial {
  unsigned long long count;
  unsigned long long sum_lo, sum_hi;
  unsigned long long sq_lo, sq_hi;
  uint32_t max_bits;  // f32 bits of the largest value (positive floats order as u32)
  uint32_t flags;     // bit 0: fixed-point range exceeded; bit 1: value outside top16 range
};

struct StatsResult {
  uint64_t count = 0;
  double sum = 0.0, sum_of_squares = 0.0, maximum = 0.0;
  uint32_t flags = 0;
};

// Radix select of the order statistics behind the five quantiles of
// nonzero_stats: rank[2q] and rank[2q+1] are the lower and upper ranks of
// quantile q (src/metrics.rs:271-291), computed on the device from the count.
constexpr int kMaxRanks = 10;
struct SelectState {
  uint32_t prefix[kMaxRanks];          // value bits found so far (the value at the end)
  unsigned long long rank[kMaxRanks];  // rank within the values sharing prefix
  uint32_t nranks;
  uint32_t pad;
};

// Tunables of the kernels (chosen per shape class by the autotuner).
// A Y-pass tile holds up to y_rows rows of 64 columns in LDS (saturated 16-bit
// values, two rows per word): the whole column when H <= y_rows, otherwise a
// chunk of y_rows - 2*y_halo rows (rounded down to even) plus a halo on each
// side (candidates beyond the halo are read from HBM).  The Z pass is the
// same with z_rows lines of 32 columns.
// fill_cap bounds the candidates the fill holds in LDS per round (at most
// kFillCap); a tile with more candidates runs several rounds.
struct Tuning {
  uint32_t y_rows = 512;
  uint32_t y_halo = 64;
  uint32_t z_rows = 512;
  uint32_t z_halo = 64;
  uint32_t fill_cap = 0;  // 0: the kernel's capacity
  // Tile order of the ridge, fill and clean-up kernels (bits 0, 1, 2): set =
  // 4x4x4 groups of tiles, clear = x-fastest.
  uint32_t tile_order = 0;
  // Tiles per fill block along x (1: 512 threads per 8^3 tile; 2: 1024
  // threads per 16x8x8 pair sharing one candidate gathering).  One value per
  // plan: the first item's.
  uint32_t fill_tiles = 1;
  // Most clean-up blocks per item (a multiple of 8 up to kStatsBlocks; 0:
  // kStatsBlocks).  384 makes the two phases of a volume one resident wave of
  // blocks (256 CUs x 3 blocks of 512 threads), so no block starts after
  // others have finished: for 512^3 volumes (4096 footprints per item) that
  // beats 1024 blocks in 2.67 rounds; for 1024x1024x1536 (49152 per item)
  // 1024 blocks are faster.  Set in the table by measurement, not swept.
  uint32_t cleanup_blocks = 0;
  // 1: the fill tests a candidate with one 4 x u8 dot product when the
  // item's radii are at most 64; 0: the packed-position test at every radius
  // (both forms give identical maps; 0 is a test override, not tuned).
  uint32_t fill_dot = 1;
  // Test overrides of the fill's fast paths (1: as the kernel chooses; 0:
  // never): list chunks looked up in the LDS chunk descriptors (else a
  // binary search over the chunk offsets), candidates staged in LDS (else
  // every round re-reads the lists).  Results never depend on them.
  uint32_t fill_chunk_map = 1;
  uint32_t fill_stage = 1;
};

constexpr int kStatsBlocks = 1024;  // most clean-up blocks of one item
// Top-16-bit histogram of thickness values: f32 bit patterns 0x4000xxxx to
// 0x47FFxxxx, i.e. values in [2, 65536).  Every diameter of the clean-up is
// at least 2 and below 2*sqrt(3)*16384.
constexpr uint32_t kTopBase = 0x4000u;
constexpr int kTopBins = 2048;
constexpr int kEdtXBlocks = 4096;
constexpr uint32_t kMinZChunk = 16;  // smallest z extent of a split Z-pass tile
constexpr uint32_t kXcds = 8;        // XCDs of the device (round-robin block dispatch)

// Item flags.
constexpr uint32_t kItemThickness = 1u;  // run the local-thickness stages
constexpr uint32_t kItemQuantiles = 2u;  // run the quantile selection (needs the map)
constexpr uint32_t kItemDiag = 4u;       // fill work counters
constexpr uint32_t kItemMap = 8u;        // the clean-up writes the thickness map
constexpr uint32_t kItemPackedFill = 16u;  // fill: packed-position test at every radius
constexpr uint32_t kItemNoChunkMap = 32u;  // fill: chunks found by binary search (test override)
constexpr uint32_t kItemNoStage = 64u;     // fill: candidates never staged (test override)

struct TileMeta;

// Everything the kernels need of one item (device memory, read with scalar
// loads).  Block counts are per stage; the Plan's start tables turn a block
// index into (item, block of the item).
struct ItemDesc {
  const uint8_t* data;
  const uint8_t* roi;
  uint16_t* dx;        // X pass output (N u16)
  uint32_t* h;         // Y pass output, then the propagated map (N u32)
  uint32_t* d2;        // squared EDT (N u32)
  uint32_t* lists;     // per-tile sorted ridge lists (512 per tile)
  TileMeta* meta;      // per-tile list header
  unsigned long long* masks;  // per tile: 8 selected + 8 target words
  uint32_t* overflow;  // [count, tile ids...]
  uint32_t* bigtiles;  // [count, tile ids...]: tiles holding ridges of r^2 > kGatherMaxR2
  uint32_t* tmpl;      // 3 x (bound+1) templates
  float* diam;         // (bound+1) diameters 2*sqrt(m)
  PhaseStatus* status;
  BlockPartial* xpart;     // one per X-pass block
  uint32_t* zmax;          // one per Z-pass block
  StatsPartial* partials;  // one per clean-up block
  StatsPartial* stats;     // final record
  uint32_t* top16;         // top-16-bit value histogram (quantiles)
  float* map;              // thickness map (kItemMap)
  SelectState* sel;
  unsigned long long* qhist;  // kMaxRanks x 256
  // the first digit pass's candidates: block b of the pass keeps those of its
  // range [b qper, (b + 1) qper) of the map at the same offsets of cand, and
  // their count in qcnt[b] (cand is the propagated map's buffer, dead by then)
  uint32_t* cand;
  uint32_t* qcnt;
  uint64_t qper;
  unsigned long long* diag;   // 8 x 16 words per fill block (kItemDiag)
  uint64_t n;
  uint32_t W, H, D, TX, TY, TZ, ntiles, bound;
  uint32_t want_nonzero, flags;
  uint32_t xblocks;
  uint32_t ygx, ygy, ygz, ychunk, yhalo;
  uint32_t zgx, zgy, zgz, zchunk, zhalo;
  uint32_t lblocks, fill_cap, cblocks, qblocks;
  uint32_t tile_order;  // Tuning::tile_order
  uint32_t r2_limit;    // ridges of r^2 >= r2_limit set status error bit 0
  // invariant-divisor constants (d, m, s1, s2) of TX, TY, (TX + 3) / 4 and
  // (TY + 3) / 4 for the tile kernels' position -> tile mapping
  uint32_t divs[4][4];
  float scale;
  int lsb;
};

// Stages with per-item block counts (start tables).
enum Stage : int {
  kStEdtX = 0,
  kStEdtY,
  kStEdtZ,
  kStLuts,
  kStRidges,    // ridges (one block per tile, padded to a multiple of 8)
  kStTiles,     // fill gather (one block per 1 or 2 tiles, padded to a multiple of 8)
  kStOverflow,  // fill overflow and scatter fallback
  kStCleanup,
  kStQHist,
  kNumStages
};

// Host description of one item.
struct ItemSpec {
  PhaseSource src;
  Dims dims;
  Tuning tune;
  uint32_t* d2_out = nullptr;  // squared EDT destination (null: internal)
  float* map_out = nullptr;    // thickness map destination (null: internal if needed)
  bool thickness = true;
  bool quantiles = false;
  bool want_map = false;       // keep the map (internal when map_out is null)
  bool keep_d2 = true;         // false: an internal map may reuse the squared EDT buffer
  bool diag = false;
  float scale = 1.0f;          // voxel_spacing[0] as f32 (src/metrics.rs:114,120)
  int lsb = 0;
};

// What comes back of one item.
struct ItemResult {
  PhaseStatus status;
  StatsPartial stats;
  SelectState sel;
};

// A growable device arena.
class Arena {
 public:
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena();
  void reserve(size_t bytes);
  void* base() const { return base_; }
  size_t capacity() const { return cap_; }

 private:
  void* base_ = nullptr;
  size_t cap_ = 0;
};

// Device buffers, descriptors and launch geometry of a list of items.
class Plan {
 public:
  Plan() = default;
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;
  ~Plan();
  // Lays out and (re)allocates the buffers of these items and uploads the
  // descriptors (synchronously on s).  With ext, the buffers are laid out in
  // that device region instead (at most ext_cap bytes; nothing is allocated
  // for them).  With ext_meta as well, the descriptors, start tables and
  // results live at the end of that region too and results come back through
  // pageable host memory: the plan then makes no HIP allocation at all.
  void build(const std::vector<ItemSpec>& specs, hipStream_t s, void* ext = nullptr,
             size_t ext_cap = 0, bool ext_meta = false);
  // Enqueues every stage for every item on s.  No allocation, no host
  // synchronisation, no memset nodes: safe under stream capture.
  void launch(hipStream_t s) const;
  // Gathers the per-item results on the device (part of launch) and copies
  // them to the host; waits for s.
  void fetch(hipStream_t s);
  size_t size() const { return desc_.size(); }
  const ItemDesc& desc(size_t i) const { return desc_[i]; }
  const ItemDesc* device_desc() const { return d_desc_; }
  const ItemResult& result(size_t i) const { return results_[i]; }
  size_t device_bytes() const { return bytes_; }
  // Device bytes a plan of these items needs (for admission control).
  static size_t bytes_for(const std::vector<ItemSpec>& specs);
  // Device bytes of the descriptors, start tables and results of nitems items.
  static size_t meta_bytes_for(size_t nitems);
  // Allocates the descriptor and result space of plans of up to nitems items
  // now, so that later builds of that size allocate nothing.
  void reserve_items(size_t nitems);

 private:
  std::vector<ItemDesc> desc_;
  std::vector<uint32_t> starts_;  // kNumStages x (nitems + 1)
  uint32_t total_[kNumStages] = {};
  uint32_t y_lds_ = 0, z_lds_ = 0;
  uint32_t fill_tiles_ = 1;  // fill variant of the whole plan
  uint32_t any_ = 0;  // union of item flags
  bool ext_meta_ = false;  // descriptors and results in the caller's region
  size_t bytes_ = 0;
  Arena mem_;                         // item buffers
  Arena meta_;                        // descriptors, start tables, results
  ItemDesc* d_desc_ = nullptr;
  uint32_t* d_starts_ = nullptr;
  ItemResult* d_results_ = nullptr;
  ItemResult* h_results_ = nullptr;  // pinned
  size_t h_results_cap_ = 0;
  std::vector<ItemResult> results_;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  std::mutex& mutex() { return mu_; }
  Plan& plain_plan() { return plain_; }  // reused by the one-volume calls
  void* scratch_bytes(size_t bytes, int slot);  // staging buffers (H2D/D2H)
  // phase_inside (src/volume.rs:84-105) as a device map.
  void run_phase_inside(const uint8_t* bone, const uint8_t* roi, uint64_t n, bool select_bone,
                        uint8_t* out, hipStream_t s);
  // Diagnostic products of a one-item plan that has run: the sphere-fill work
  // into its status, ridge flags from the tile lists (and the ridge count into
  // its status) and the propagated map.
  void stage_outputs(const Plan& p, uint8_t* ridge_flags_out, uint32_t* propagated_out,
                     hipStream_t s);

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  std::mutex mu_;
  Plan plain_;
  Arena stage_[4];  // slots 0-2: staging and scratch volumes; 3: small tables
};

// Per-kernel timing with HIP events recorded on the launching stream (off by
// default).  Used by the benchmark to measure kernel durations live.
struct KernelTime {
  char name[48];
  uint64_t launches;
  double total_ms;
};
void profile_enable(bool on);
bool profile_enabled();
// Whether launches of the kernel named name are timed: profiling on and the
// name in the set given to profile_only (every name when the set is empty).
bool profile_enabled_for(const char* name);
void profile_only(const char* names);  // comma-separated names; null or "" = all
hipEvent_t profile_begin();
void profile_end(const char* name, hipEvent_t start, hipStream_t s);
int profile_read(KernelTime* out, int max);  // synchronises, aggregates, resets
// Profiling is suspended while a stream is being captured into a graph.
void profile_suspend(bool on);

// Launch a kernel, bracketed by profiling events when profiling is on.
#define BM_LAUNCH(name, kernel, grid, block, lds, stream, ...)           \
  do {                                                                   \
    hipEvent_t bm_ev0_ = ::bm::profile_enabled_for(name) ? ::bm::profile_begin() \
                                                 : nullptr;              \
    if (bm_ev0_) BM_HIP(hipEventRecord(bm_ev0_, stream));                \
    kernel<<<grid, block, lds, stream>>>(__VA_ARGS__);                   \
    if (bm_ev0_) ::bm::profile_end(name, bm_ev0_, stream);               \
  } while (0)

// Engine of a device (created on first use; process lifetime).
Engine& engine_for_device(int device);
int current_device();

// Host-side conversion of the exact accumulator into f64 sums.
StatsResult finish_stats(const StatsPartial& p, int lsb_exp);
int stats_lsb_exp(float scale_f32);

// Tunables per shape class, from the autotuner's JSON cache (defaults when
// the cache has no entry): see bm_tune.cpp.
Tuning tuned_for(const Dims& dims);
std::string shape_class(const Dims& dims);

// The five quantiles of nonzero_stats from a finished selection.
struct QuantileResult {
  double median = 0, p10 = 0, p25 = 0, p75 = 0, p90 = 0;
};
QuantileResult quantiles_from(const SelectState& sel, uint64_t count, float scale_f32);

}  // namespace bm
