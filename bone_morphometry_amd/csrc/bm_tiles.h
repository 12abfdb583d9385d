// bm_tiles.h -- layout of the 8x8x8 tile products of local thickness.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bm {

// Per tile: number of ridges, their largest r^2, and prefix counts: pref[j] is
// an upper bound on the number of list entries with r^2 >= kPrefR2(j), i.e. the
// number of entries whose bucket key r^2 >> shift is at least kPrefR2(j) >> shift.
// Every list is sorted by decreasing bucket key; shift is the same for all
// tiles of a volume (from the volume's largest squared distance).
struct TileMeta {
  uint32_t count;
  uint32_t maxr2;
  uint16_t pref[16];
};
// (the fill reads count and maxr2 as one 8-byte word)
static_assert(sizeof(TileMeta) == 40 && offsetof(TileMeta, maxr2) == 4, "TileMeta layout");

constexpr int kBucketBits = 9;
constexpr int kBuckets = 1 << kBucketBits;  // radius buckets of the counting sorts
constexpr int kFillCap = 1024;              // LDS candidates per round of the fill
constexpr int kStageCap = 2048;             // candidates of a tile staged in LDS
constexpr int kChunkMap = 2048;             // 64-entry list chunks mapped to their tile in LDS
constexpr int kMaxNeighbours = 512;         // tiles in reach kept per tile
// Ridges of radius above kGatherMaxRadius (r^2 > kGatherMaxR2) are left out
// of the gather and scattered afterwards (atomic maxima); the gather's reach
// and candidate keys assume radii up to this bound.
constexpr uint32_t kGatherMaxRadius = 128;
constexpr uint32_t kGatherMaxR2 = kGatherMaxRadius * kGatherMaxRadius;
// A list entry is r^2 << 9 | local index: r^2 must stay below 2^23.
constexpr uint32_t kListR2Limit = 1u << 23;
constexpr int kDiamLds = 2048;              // diameters the clean-up keeps in LDS
constexpr uint32_t kOverflowBlocks = 64;    // blocks per item of the fill fallbacks and the scatter

// Radius thresholds of the prefix counts: (2j + 1)^2.
inline __host__ __device__ uint32_t kPrefR2(int j) { return (uint32_t)(2 * j + 1) * (2 * j + 1); }

inline uint32_t tiles_of(uint32_t n) { return (n + 7) / 8; }

// Work items of the ridge pass: 64 x 8 (x, y) footprints (8 tiles along x)
// over chunks of kRidgeZ planes (whole tile layers).
constexpr uint32_t kRidgeZ = 64;
inline __host__ __device__ uint32_t ridge_work_items(uint32_t W, uint32_t H, uint32_t D) {
  return ((W + 63) / 64) * ((H + 7) / 8) * ((D + kRidgeZ - 1) / kRidgeZ);
}

// Work items of the clean-up: 64 x 8 (x, y) footprints over chunks of
// kCleanZ planes.
constexpr uint32_t kCleanZ = 64;
inline __host__ __device__ uint32_t cleanup_work_items(uint32_t W, uint32_t H, uint32_t D) {
  return ((W + 63) / 64) * ((H + 7) / 8) * ((D + kCleanZ - 1) / kCleanZ);
}

}  // namespace bm
