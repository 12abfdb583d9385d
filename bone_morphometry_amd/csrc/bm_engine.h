// bm_engine.h -- device engine for the exact squared EDT and local thickness.
//
// Everything here is stream-ordered device work on one GPU.  The host-facing
// mirror of the reference crate (BinaryVolume, Phase, euclidean_distance_
// transform_squared, local_thickness, analyze) lives in bm_api.cpp and calls
// this engine; the C-ABI in include/bm.h is built on top of both.
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

// Per-phase device status, written by the kernels of one phase.
struct PhaseStatus {
  unsigned long long selected;     // voxels of the phase (X pass)
  unsigned long long ridge_count;  // distance ridges found
  unsigned long long fill_visits;  // voxel-candidate tests of the fill
  uint32_t has_site;               // any non-selected voxel (X pass)
  uint32_t max_d2;                 // maximum squared distance (Z pass)
  uint32_t error;                  // bit 0: r^2 too large for the tile lists
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
  uint32_t flags;
};

struct StatsResult {
  uint64_t count = 0;
  double sum = 0.0, sum_of_squares = 0.0, maximum = 0.0;
  uint32_t flags = 0;
};

// Tunables of the kernels (chosen by the autotuner for the batched path).
// A Y-pass tile holds up to y_rows rows of 64 columns (u16) in LDS: the whole
// column when H <= y_rows, otherwise a chunk of y_rows - 2*y_halo rows plus a
// halo on each side (candidates beyond the halo are read from HBM).  The Z
// pass is the same with z_rows lines of 32 columns (u32).
// fill_cap bounds the candidates the fill holds in LDS per round (at most
// kFillCap); a tile with more candidates runs several rounds.
struct Tuning {
  uint32_t y_rows = 512;
  uint32_t y_halo = 64;
  uint32_t z_rows = 512;
  uint32_t z_halo = 64;
  uint32_t fill_cap = 0;  // 0: the kernel's capacity
};

// Device buffers for one phase of one volume.
struct TileMeta;
struct PhaseBuffers {
  uint16_t* dx = nullptr;       // X pass output (N u16)
  uint32_t* h = nullptr;        // Y pass output, then the propagated map (N u32)
  uint32_t* d2 = nullptr;       // squared EDT (N u32)
  uint32_t* lists = nullptr;    // per-tile sorted ridge lists (512 per tile)
  TileMeta* meta = nullptr;     // per-tile list header
  unsigned long long* masks = nullptr;  // per tile: 8 selected + 8 target words
  uint32_t* overflow = nullptr; // [count, tile ids...]
  uint32_t tiles[3] = {0, 0, 0};
  uint32_t* tmpl = nullptr;     // 3 x (bound+1) templates
  float* diam = nullptr;        // (bound+1) diameters 2*sqrt(m)
  PhaseStatus* status = nullptr;
  BlockPartial* xpart = nullptr;     // kEdtXBlocks
  uint32_t* zmax = nullptr;          // one per Z-pass block
  StatsPartial* partials = nullptr;  // kStatsBlocks
  StatsPartial* stats = nullptr;     // final record
  uint32_t* top16 = nullptr;         // top-16-bit value histogram (quantiles)
  uint64_t bound = 0;
};

constexpr int kStatsBlocks = 1024;
// Top-16-bit histogram of thickness values: f32 bit patterns 0x4000xxxx to
// 0x47FFxxxx, i.e. values in [2, 65536).  Every diameter of the clean-up is
// at least 2 and below 2*sqrt(3)*16384.
constexpr uint32_t kTopBase = 0x4000u;
constexpr int kTopBins = 2048;
constexpr int kEdtXBlocks = 4096;
constexpr uint32_t kMinZChunk = 64;  // smallest z extent of a Z-pass tile

// A growable device arena, freed by the engine.
class Arena {
 public:
  ~Arena();
  void reserve(size_t bytes);
  void* base() const { return base_; }
  size_t capacity() const { return cap_; }

 private:
  void* base_ = nullptr;
  size_t cap_ = 0;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  std::mutex& mutex() { return mu_; }

  // Workspace for a volume of these dims (both phases share it sequentially).
  PhaseBuffers buffers_for(const Dims& dims);
  void* scratch_bytes(size_t bytes, int slot);  // extra staging buffers (H2D/D2H)

  // Squared EDT into d2 (device pointer, may be buffers.d2).
  void run_edt(const PhaseSource& src, const Dims& dims, const PhaseBuffers& b, uint32_t* d2,
               hipStream_t s, const Tuning& t = Tuning());
  // phase_inside (src/volume.rs:84-105) as a device map.
  void run_phase_inside(const uint8_t* bone, const uint8_t* roi, uint64_t n, bool select_bone,
                        uint8_t* out, hipStream_t s);
  // Full local thickness from an EDT already in b.d2: ridges, fill, clean-up,
  // phase mask, statistics.  map_out may be null (stats only).
  void run_thickness(const Dims& dims, const PhaseBuffers& b, float* map_out, float scale_f32,
                     int lsb_exp, hipStream_t s, uint8_t* ridge_flags_out = nullptr,
                     uint32_t* propagated_out = nullptr, const Tuning& t = Tuning());

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  std::mutex mu_;
  Arena work_;
  Arena stage_[3];
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
hipEvent_t profile_begin();
void profile_end(const char* name, hipEvent_t start, hipStream_t s);
int profile_read(KernelTime* out, int max);  // synchronises, aggregates, resets

// Launch a kernel, bracketed by profiling events when profiling is on.
#define BM_LAUNCH(name, kernel, grid, block, lds, stream, ...)           \
  do {                                                                   \
    hipEvent_t bm_ev0_ = ::bm::profile_enabled() ? ::bm::profile_begin() \
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

}  // namespace bm
