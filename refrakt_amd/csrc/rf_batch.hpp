// Batched rendering: device-side frame table and kernel arguments.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "rf_device.hpp"
#include "rf_sort.hpp"

namespace rfk {

struct BatchFrame {
  float* bins;            // W*H RGBA, rows bottom-up
  uint32_t W, H;
  uint32_t item_begin;    // first work item of the frame (frames are sorted by it)
  uint32_t n_items;
  float ss[6];            // screen affine
  uint32_t goff;          // first bin of the frame in the batch histogram (a multiple of 2^12)
  uint32_t pad;
};
static_assert(sizeof(BatchFrame) == 56, "BatchFrame layout");

struct BatchArgs {
  const BatchFrame* frames;
  const rfd::Block* blocks;       // one per frame
  const float4* palettes;         // 256 per frame
  const uint16_t* perms;          // NPERM rows x (64*waves)
  unsigned long long* drawn;      // per frame: binned samples
  uint32_t* remaining;            // per frame: items whose samples are not yet in the histogram
  unsigned long long* done_ts;    // per frame: wall clock when the last item was delivered
  uint64_t seed;
  uint32_t launch;
  uint32_t n_frames;
  uint32_t item0, n_items;        // the launch's work items: item0 .. item0 + n_items - 1
  uint32_t fuse, n_iter;          // unbinned and binned iterations per particle
  // RF_HIST_SORT: a launch (slice) stages records for the bins of its span,
  // numbered from span_base in the batch histogram; span_bins points there
  // (records beyond a bucket's capacity are added straight to it).
  uint32_t span_base;
  float* span_bins;
  SortArgs sort;
};

#ifndef __HIPCC_RTC__
// Caches the resident-workgroup count of a kernel variant (call before a capture).
hipError_t batch_prepare(int waves, int hist_mode);
hipError_t launch_batch(const BatchArgs& a, int waves, int hist_mode, uint32_t max_grid, hipStream_t s);
hipError_t launch_batch_init(const BatchArgs& a, float4* hist, uint64_t n_bins, const uint32_t* items,
                             uint32_t* tiles_done, unsigned long long* t0, hipStream_t s);
#endif

}  // namespace rfk
