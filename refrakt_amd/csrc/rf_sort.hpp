// Sort-based histogram accumulation (RF_HIST_SORT): shared definitions.
//
// Instead of one memory-side atomic per sample, the chaos-game kernel writes
// 4-byte records into coarse buckets of 2^19 bins, a second kernel splits
// each bucket into 128 tiles of 2^12 bins, and a third accumulates every tile
// in LDS and adds it to the histogram with one plain read-modify-write per
// touched bin.  Records carry no weight, so the mode needs unit opacities;
// the LDS sums are fixed point, so palette values must lie in [0, 1].
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace rfk {

constexpr int SORT_COARSE_SHIFT = 19;                                        // bins per bucket: 2^19
constexpr int SORT_TILE_SHIFT = 12;                                          // bins per tile: 2^12
constexpr int SORT_TILES_PER_BUCKET = 1 << (SORT_COARSE_SHIFT - SORT_TILE_SHIFT);  // 128
constexpr int SORT_MAX_BUCKETS = 256;                                        // 2^27 bins
constexpr int SORT_STAGE_SLOTS = 4096;                                       // samples staged in LDS per flush
constexpr int SORT_CHUNK = 16384;                                            // records per rebucketing workgroup
// Tile sums add red | green << 32 and blue | count << 32 per record with
// 64-bit LDS atomics, colours in fixed point.  A tile of at most SORT_SEG16
// records is summed by one workgroup at 2^16 - 1 per unit (a field then
// holds at most 65532 * 65535 < 2^32, so no carry crosses into the next); a
// larger tile is split into segments of at most SORT_SEG15 records, each
// summed by its own workgroup at 2^15 - 1 per unit (131068 * 32767 < 2^32).
// Each sample's colour is rounded by at most 1 / (2 * 65535) or
// 1 / (2 * 32767), about 2^-17 and 2^-16.
constexpr int SORT_SEG16 = 65532;
constexpr int SORT_SEG15 = 131068;
constexpr uint32_t SORT_NONE = 0xffffffffu;

struct SortArgs {
  uint32_t* cbuf;          // bucket records: ((bin mod 2^19) << 8) | palette index
  uint32_t* tbuf;          // tile records:   ((bin mod 2^12) << 8) | palette index
  uint32_t* bcur;          // per bucket: records offered (may exceed the capacity)
  uint32_t* tcur;          // per tile: records offered
  const uint32_t* boff;    // per bucket: first record in cbuf
  const uint32_t* bcap;    // per bucket: capacity
  const uint32_t* toff;    // per tile: first record in tbuf
  const uint32_t* tcap;    // per tile: capacity
  const uint32_t* chunks;  // rebucketing work: (bucket << 16) | chunk index, one per workgroup
  const uint32_t* segs;    // tile work: (tile << 16) | segment index, one per workgroup
  uint32_t n_buckets;
  uint32_t n_tiles;        // n_buckets * SORT_TILES_PER_BUCKET
  uint32_t n_chunks, n_segs;
  // Batched rendering: per tile, the offset (in entries) of its frame's
  // palette in the palette array the record passes get; null when one
  // palette serves every tile.
  const uint32_t* tile_pal;
};

#ifndef __HIPCC_RTC__
// Clears the cursors and, when counter is given, counter_words words of it.
hipError_t launch_sort_reset(const SortArgs& a, hipStream_t s, unsigned long long* counter = nullptr,
                             uint32_t counter_words = 0);
// Copies counter_words 64-bit counters, then nb bucket and nt tile cursors
// (either array may be null) into out, 32-bit words, in one launch; out may
// be the device address of pinned host memory.
hipError_t launch_draw_readback(const unsigned long long* counter, uint32_t counter_words, const uint32_t* bcur,
                                uint32_t nb, const uint32_t* tcur, uint32_t nt, uint32_t* out, hipStream_t s);
hipError_t launch_sort_rebucket(const SortArgs& a, float* bins, const float4* palette, hipStream_t s);
hipError_t launch_sort_tiles(const SortArgs& a, float* bins, uint32_t n_bins, const float4* palette, hipStream_t s);
// mass[t] = sum of the count channel over tile t of a histogram (t < n_tiles).
hipError_t launch_sort_tile_mass(const float4* bins, uint32_t n_bins, uint32_t n_tiles, float* mass, hipStream_t s);
#endif

}  // namespace rfk
