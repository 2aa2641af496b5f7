// Batched rendering: many frames (each its own flame, size and sample count)
// in one persistent chaos-game launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rf_k_batch_body.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

template <int WAVES, int HM, int LOGK>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(HM == 5 ? 6 : 1)))
void k_batch(BatchArgs g, const Block* __restrict__ blocks, const BatchFrame* __restrict__ frames) {
  k_batch_body<WAVES, HM, LOGK, InterpProg>(g, blocks, frames);
}

// Zeroes every frame's histogram, the per-frame counters and DE tile
// counters, re-arms `remaining`, and stamps the batch start.  A kernel rather
// than memset nodes, so the captured graph replays reliably.
__global__ void k_batch_init(BatchArgs g, float4* __restrict__ hist, uint64_t n_bins,
                             const uint32_t* __restrict__ items, uint32_t* __restrict__ tiles_done,
                             unsigned long long* __restrict__ t0) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = i0; i < n_bins; i += stride) hist[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i0 < g.n_frames) {
    g.drawn[i0] = 0ull;
    g.remaining[i0] = items[i0];
    g.done_ts[i0] = 0ull;
    tiles_done[i0] = 0u;
  }
  if (i0 == 0) *t0 = wall_clock64();
}

}  // namespace

template <int W, int HM, int LOGK>
static hipError_t resident_k(int* out) {
  static int resident = 0;
  if (!resident) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_batch<W, HM, LOGK>),
                                                       64 * W, 0);
    if (e != hipSuccess) return e;
    resident = cus * (per_cu > 0 ? per_cu : 1);
  }
  *out = resident;
  return hipSuccess;
}

template <int W, int HM, int LOGK>
static hipError_t launch_batch_k(const BatchArgs& a, uint32_t max_grid, hipStream_t s) {
  int resident = 0;
  const hipError_t e = resident_k<W, HM, LOGK>(&resident);
  if (e != hipSuccess) return e;
  // LDS-table and direct modes: one workgroup per resident slot; sort
  // staging: one per item, like the fast engine's sort kernel
  uint32_t grid = HM == 5 ? a.n_items : std::min<uint32_t>(a.n_items, (uint32_t)resident);
  if (max_grid) grid = std::min(grid, max_grid);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_batch<W, HM, LOGK>), dim3(grid), dim3(64 * W), 0, s, a, a.blocks, a.frames);
  return hipGetLastError();
}

hipError_t batch_prepare(int waves, int hist_mode) {
  int r = 0;
  if (waves == 8) {
    if (hist_mode == 4) return resident_k<8, 4, 11>(&r);
    return hist_mode == 5 ? resident_k<8, 5, 0>(&r) : resident_k<8, 1, 0>(&r);
  }
  if (hist_mode == 4) return resident_k<4, 4, 10>(&r);
  return hist_mode == 5 ? resident_k<4, 5, 0>(&r) : resident_k<4, 1, 0>(&r);
}

hipError_t launch_batch(const BatchArgs& a, int waves, int hist_mode, uint32_t max_grid, hipStream_t s) {
  if (waves == 8) {
    if (hist_mode == 4) return launch_batch_k<8, 4, 11>(a, max_grid, s);
    return hist_mode == 5 ? launch_batch_k<8, 5, 0>(a, max_grid, s) : launch_batch_k<8, 1, 0>(a, max_grid, s);
  }
  if (hist_mode == 4) return launch_batch_k<4, 4, 10>(a, max_grid, s);
  return hist_mode == 5 ? launch_batch_k<4, 5, 0>(a, max_grid, s) : launch_batch_k<4, 1, 0>(a, max_grid, s);
}

hipError_t launch_batch_init(const BatchArgs& a, float4* hist, uint64_t n_bins, const uint32_t* items,
                             uint32_t* tiles_done, unsigned long long* t0, hipStream_t s) {
  const uint64_t want = (n_bins + 255) / 256;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, 8192));
  hipLaunchKernelGGL(k_batch_init, dim3(grid), dim3(256), 0, s, a, hist, n_bins, items, tiles_done, t0);
  return hipGetLastError();
}

}  // namespace rfk
