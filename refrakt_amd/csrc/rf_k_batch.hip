// Batched rendering: many frames (each its own flame, size and sample count)
// in one persistent chaos-game launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rf_batch.hpp"
#include "rf_kern_common.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

typedef const __attribute__((address_space(4))) BatchFrame CFrame;
__device__ __forceinline__ CFrame* frame_constant(const BatchFrame* p) { return (CFrame*)p; }

// Work item = 64*WAVES fresh particles of one frame: they start uniformly in
// [-1,1]^2, run `fuse` unbinned and `n_iter` binned iterations in registers
// and are discarded (a frame rendered from scratch keeps no particle state).
// Items are numbered frame by frame; a workgroup takes items blockIdx.x,
// blockIdx.x + gridDim.x, ..., so consecutive items of a workgroup mostly
// belong to the same frame.  The iteration is the fast engine's (wave-uniform
// xform per iteration, LDS permutation of the item's particles every
// iteration); the random streams are keyed by (item, lane / wave / item), so
// results do not depend on the grid.
//
// HM 4 keeps an LDS bin table per workgroup for the frame it is working on
// and flushes it when the frame changes and at the end.  `remaining[f]`
// counts the frame's items whose samples are not yet in its histogram; the
// workgroup that delivers the last one stamps done_ts[f].
template <int WAVES, int HM, int LOGK>
__global__ __launch_bounds__(64 * WAVES) void k_batch(BatchArgs g, const Block* __restrict__ blocks,
                                                      const BatchFrame* __restrict__ frames) {
  constexpr int N = 64 * WAVES;
  __shared__ float s_x[2][N], s_y[2][N], s_c[2][N];
  __shared__ float4 s_pal[256];
  __shared__ uint16_t s_perm[NPERM * N];
  __shared__ LdsBins<LOGK> s_bins;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < (uint32_t)(NPERM * N); i += N) s_perm[i] = g.perms[i];
  if constexpr (HM == 4) lds_bins_clear(s_bins);

  int cur = -1;              // frame whose palette (and LDS table) is loaded
  uint32_t pend = 0;         // items of `cur` not yet delivered
  unsigned long long drawn = 0;
  uint32_t hits = 0;

  // Makes the samples of the pending items of frame `f` visible and counts them.
  auto deliver = [&](int f) {
    CFrame* F = frame_constant(frames + f);
    __syncthreads();
    if constexpr (HM == 4) {
      flush_lds(s_bins, F->bins);
      __syncthreads();
      lds_bins_clear(s_bins);
    }
    const unsigned long long wd = drawn;  // already wave-uniform (ballot counts)
    if ((tid & (WAVE - 1)) == 0 && wd) atomicAdd(g.drawn + f, wd);
    drawn = 0;
    __syncthreads();
    if (tid == 0) {
      __threadfence();
      if (atomicSub(g.remaining + f, pend) == pend) {
        const unsigned long long t = wall_clock64();
        __hip_atomic_store(g.done_ts + f, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    pend = 0;
  };

  for (uint32_t item = blockIdx.x; item < g.n_items; item += gridDim.x) {
    // frame of the item (frames are sorted by item_begin; the search is uniform)
    uint32_t f = 0;
    while (f + 1 < g.n_frames && frame_constant(frames + f + 1)->item_begin <= item) ++f;
    f = __builtin_amdgcn_readfirstlane(f);
    if ((int)f != cur) {
      if (cur >= 0) deliver(cur);
      __syncthreads();
      for (uint32_t i = tid; i < 256u; i += N) s_pal[i] = g.palettes[(size_t)f * 256 + i];
      if constexpr (HM == 4) lds_bins_palette(s_bins, g.palettes + (size_t)f * 256);
      __syncthreads();
      cur = (int)f;
    }
    CBlock* blk = as_constant(blocks + f);
    CFrame* F = frame_constant(frames + f);
    const float ss[6] = {F->ss[0], F->ss[1], F->ss[2], F->ss[3], F->ss[4], F->ss[5]};
    const uint32_t W = F->W, H = F->H;
    float* bins = F->bins;

    const uint32_t pid = item * (uint32_t)N + tid;
    Jsf32 lrng = jsf_from_philox(g.seed, pid, g.launch, 0u);
    const uint32_t wid = __builtin_amdgcn_readfirstlane(item * WAVES + (tid >> 6));
    Jsf32 wrng = jsf_from_philox(g.seed, wid, g.launch, 1u);
    Jsf32 grng = jsf_from_philox(g.seed, item, g.launch, 2u);
    float x = lrng.randf() * 2.0f - 1.0f;
    float y = lrng.randf() * 2.0f - 1.0f;
    float c = lrng.randf();

    const uint32_t total = g.fuse + g.n_iter;
    for (uint32_t it = 0; it < total; ++it) {
      const int xid = select_xform_raw(blk, wrng.next_a());
      XOut o = apply_xform(blk, xid, x, y, c, false, lrng);
      if (rf_badval(o.x) || rf_badval(o.y)) {
        o.x = lrng.randf() * 2.0f - 1.0f;
        o.y = lrng.randf() * 2.0f - 1.0f;
        o.w = 0.0f;
      }
      x = o.x; y = o.y; c = o.c;
      if (it >= g.fuse) {
        const Located L = locate(blk, o, lrng, ss, W, H);
        if constexpr (HM == 4) bin_lds(s_bins, bins, L.bidx, L.valid, L.pidx, L.w, s_pal, hits);
        else bin_sample<HM>(bins, L.bidx, L.valid, L.pidx, L.w, s_pal);
        drawn += (unsigned long long)__popcll(__ballot(L.valid));
      }
      const uint32_t row = grng.next_a() & (uint32_t)(NPERM - 1);
      const uint32_t dst = s_perm[row * N + tid];
      const uint32_t b = it & 1u;
      s_x[b][dst] = x; s_y[b][dst] = y; s_c[b][dst] = c;
      lds_barrier();
      x = s_x[b][tid]; y = s_y[b][tid]; c = s_c[b][tid];
    }
    lds_barrier();
    ++pend;
    if constexpr (HM != 4) deliver((int)f);
  }
  if (cur >= 0 && pend) deliver(cur);
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
  uint32_t grid = std::min<uint32_t>(a.n_items, (uint32_t)resident);
  if (max_grid) grid = std::min(grid, max_grid);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_batch<W, HM, LOGK>), dim3(grid), dim3(64 * W), 0, s, a, a.blocks, a.frames);
  return hipGetLastError();
}

hipError_t batch_prepare(int waves, int hist_mode) {
  int r = 0;
  const bool lds = hist_mode == 4;
  if (waves == 8) return lds ? resident_k<8, 4, 11>(&r) : resident_k<8, 1, 0>(&r);
  return lds ? resident_k<4, 4, 10>(&r) : resident_k<4, 1, 0>(&r);
}

hipError_t launch_batch(const BatchArgs& a, int waves, int hist_mode, uint32_t max_grid, hipStream_t s) {
  const bool lds = hist_mode == 4;
  if (waves == 8) return lds ? launch_batch_k<8, 4, 11>(a, max_grid, s) : launch_batch_k<8, 1, 0>(a, max_grid, s);
  return lds ? launch_batch_k<4, 4, 10>(a, max_grid, s) : launch_batch_k<4, 1, 0>(a, max_grid, s);
}

hipError_t launch_batch_init(const BatchArgs& a, float4* hist, uint64_t n_bins, const uint32_t* items,
                             uint32_t* tiles_done, unsigned long long* t0, hipStream_t s) {
  const uint64_t want = (n_bins + 255) / 256;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, 8192));
  hipLaunchKernelGGL(k_batch_init, dim3(grid), dim3(256), 0, s, a, hist, n_bins, items, tiles_done, t0);
  return hipGetLastError();
}

}  // namespace rfk
