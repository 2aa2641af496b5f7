// Batched-rendering kernel body, shared by the precompiled kernels
// (rf_k_batch.hip, flames interpreted from their Blocks) and by compiled
// flame programs (rf_jit.cpp: every frame's xform i has the same structure).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "rf_batch.hpp"
#include "rf_k_fast_body.hpp"

namespace rfk {
namespace detail {
using namespace rfd;

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
// and flushes it when the frame changes and at the end.  HM 5 stages the
// samples of an item in LDS and flushes them as bucket records of the
// launch's span (the fast engine's sort staging, sort_flush); the record
// passes that follow the launch add them to the histogram.  `remaining[f]`
// counts the frame's items whose samples have left the chaos game (are in
// its histogram, or in records); the workgroup that delivers the last one
// stamps done_ts[f].
template <int WAVES, int HM, int LOGK, class Prog>
__device__ __forceinline__ void k_batch_body(const BatchArgs& g, const Block* __restrict__ blocks,
                                             const BatchFrame* __restrict__ frames) {
  constexpr int N = 64 * WAVES;
  __shared__ float s_x[2][N], s_y[2][N], s_c[2][N];
  __shared__ float4 s_pal[HM == 5 ? 1 : 256];
  __shared__ __attribute__((aligned(16))) uint16_t s_perm[NPERM * N];
  __shared__ LdsBins<LOGK> s_bins;
  __shared__ SortStage<N, HM == 5> s_sort;
  const uint32_t tid = threadIdx.x;
  // 16-byte copies (the table is 16-byte aligned in device memory and in LDS)
  static_assert((NPERM * N) % 8 == 0, "whole 16-byte pieces");
  for (uint32_t i = tid; i < (uint32_t)(NPERM * N / 8); i += N)
    reinterpret_cast<uint4*>(s_perm)[i] = reinterpret_cast<const uint4*>(g.perms)[i];
  if constexpr (HM == 4) lds_bins_clear(s_bins);
  if constexpr (HM == 5) {
    for (uint32_t b = tid; b < g.sort.n_buckets; b += N) {
      s_sort.boff[b] = g.sort.boff[b];
      s_sort.bcap[b] = g.sort.bcap[b];
      s_sort.cnt[b] = 0u;
    }
  }

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

  const uint32_t item_end = g.item0 + g.n_items;
  for (uint32_t item = g.item0 + blockIdx.x; item < item_end; item += gridDim.x) {
    // frame of the item (frames are sorted by item_begin; the search is uniform)
    uint32_t f = 0;
    while (f + 1 < g.n_frames && frame_constant(frames + f + 1)->item_begin <= item) ++f;
    f = __builtin_amdgcn_readfirstlane(f);
    if ((int)f != cur) {
      if (cur >= 0) deliver(cur);
      __syncthreads();
      if constexpr (HM != 5)
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
      const int xid = Prog::select(blk, wrng.next_a());
      XOut o = Prog::apply(blk, xid, x, y, c, false, lrng);
      if (rf_badval(o.x) || rf_badval(o.y)) {
        o.x = lrng.randf() * 2.0f - 1.0f;
        o.y = lrng.randf() * 2.0f - 1.0f;
        o.w = 0.0f;
      }
      x = o.x; y = o.y; c = o.c;
      if (it >= g.fuse) {
        const Located L = locate<Prog>(blk, o, lrng, ss, W, H);
        if constexpr (HM == 4) {
          bin_lds(s_bins, bins, L.bidx, L.valid, L.pidx, L.w, s_pal, hits);
        } else if constexpr (HM == 5) {
          const uint32_t slot = ((it - g.fuse) % sort_stage_iters<N>()) * N + tid;
          s_sort.bin[slot] = L.valid ? F->goff + L.bidx - g.span_base : SORT_NONE;
          s_sort.pid[slot] = (uint8_t)L.pidx;
        } else {
          bin_sample<HM>(bins, L.bidx, L.valid, L.pidx, L.w, s_pal);
        }
        drawn += (unsigned long long)__popcll(__builtin_amdgcn_ballot_w64(L.valid));
      }
      const uint32_t row = grng.next_a() & (uint32_t)(NPERM - 1);
      const uint32_t dst = s_perm[row * N + tid];
      const uint32_t b = it & 1u;
      s_x[b][dst] = x; s_y[b][dst] = y; s_c[b][dst] = c;
      lds_barrier();
      x = s_x[b][tid]; y = s_y[b][tid]; c = s_c[b][tid];
      if constexpr (HM == 5) {
        // the exchange barrier above also orders the staging writes before the flush
        constexpr uint32_t SI = sort_stage_iters<N>();
        if (it >= g.fuse && (((it - g.fuse) % SI) == SI - 1 || it + 1 == total))
          sort_flush(s_sort, ((it - g.fuse) % SI) + 1, g.sort, g.span_bins, g.palettes + (size_t)f * 256);
      }
    }
    lds_barrier();
    ++pend;
    if constexpr (HM != 4) deliver((int)f);
  }
  if (cur >= 0 && pend) deliver(cur);
}

}  // namespace detail
}  // namespace rfk
