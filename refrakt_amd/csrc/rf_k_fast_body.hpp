// Fast engine kernel body, shared by the precompiled kernels (rf_k_fast.hip,
// flames interpreted from their Block) and by compiled flame programs
// (rf_jit.cpp, the same body instantiated with a flame's structure fixed).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "rf_kern_common.hpp"

namespace rfk {
namespace detail {
using namespace rfd;

// ------------------------------------------------------------ fast engine
// Persistent chaos game.  A batch is 64*WAVES particles; a workgroup takes
// batches blockIdx.x, blockIdx.x + gridDim.x, ... and runs each for all
// n_iter iterations with the particles in registers.  Per iteration each wave
// draws one xform from its own wave-uniform stream (so the variation switch
// never diverges), applies it, bins, and then the batch is permuted through
// LDS with one of NPERM random permutations chosen by a batch-uniform stream,
// which decorrelates the xform sequences of particles sharing a wave (the
// role of the reference's shuffle rows, src/shuffle_buffers.cpp).  All random
// streams are keyed by particle, wave-of-batch and batch index, never by the
// workgroup, so the results do not depend on the grid size.
//
// HM 1/2/3/5 launch one workgroup per batch; HM 4 (LDS bin cache) launches
// one workgroup per resident slot so that each workgroup's LDS table spans
// many batches before it is flushed.
// HIST_SORT staging: the samples of SI = sort_stage_iters<N>() iterations
// wait in LDS (one fixed slot per lane and iteration, no append atomics).  A
// flush ranks them within their buckets (each lane loads its staged bins
// first, then issues its rank atomics back to back and keeps the ranks in
// registers), one lane per non-empty bucket reserves the bucket's run with
// one global atomic and clears its count, and every lane stores its records
// at run start + rank.  The LDS serves one instruction's same-address
// atomics lane after lane, so the lanes of one bucket in a store instruction
// write one contiguous piece of the run, and the pieces of a run meet in L2
// (correctness needs only distinct ranks).  Two barriers per flush.  (Placing the records
// in bucket order in LDS first, so that each store instruction writes one
// contiguous span, took two more barriers and five more LDS accesses per
// record, and 5 % longer per 7680x4320 draw.)
template <int N>
constexpr int sort_stage_iters() { return SORT_STAGE_SLOTS / N; }

// Waves per SIMD the register allocation must allow: every mode but the LDS
// bin cache (HM 4, two workgroups of 8 waves per CU by its LDS) fits three
// workgroups of 8 waves per CU in LDS, six waves per SIMD, if each wave
// keeps to 80 VGPRs.
template <int HM>
constexpr int fast_waves_per_eu() { return HM == 4 ? 1 : 6; }
template <int N>
constexpr int sort_stage_slots() { return sort_stage_iters<N>() * N; }

template <int N, bool ON>
struct SortStage {
  static constexpr int S = sort_stage_slots<N>();
  uint32_t bin[S];       // staged bin index (SORT_NONE: no sample), slot = iteration * N + lane
  uint8_t pid[S];        // staged palette index
  // records per bucket in this flush (zero between flushes), then one spare
  // counter per lane of a wave that samples without a bin count in, so that
  // the rank atomics need no branch and those samples do not contend
  uint32_t cnt[SORT_MAX_BUCKETS + WAVE];
  uint2 run[SORT_MAX_BUCKETS];       // this flush's run of bucket b: first record in cbuf, records that fit
  uint32_t boff[SORT_MAX_BUCKETS];   // bucket offsets and capacities, staged once per workgroup
  uint32_t bcap[SORT_MAX_BUCKETS];
};
template <int N>
struct SortStage<N, false> {};

template <int N>
__device__ __forceinline__ void sort_flush(SortStage<N, true>& st, uint32_t n_staged, const SortArgs& a,
                                           float* __restrict__ bins, const float4* __restrict__ pal) {
  constexpr int SI = sort_stage_iters<N>();
  const uint32_t tid = threadIdx.x;
  // 1. rank this lane's staged samples within their buckets.  Every slot is
  //    loaded (the slots of a partial stage beyond n_staged hold older
  //    samples and are masked after the load), then the rank atomics are
  //    issued back to back.
  uint32_t sb[SI], rank[SI];
#pragma unroll
  for (int k = 0; k < SI; ++k) {
    const uint32_t v = st.bin[k * N + tid];
    sb[k] = (uint32_t)k < n_staged ? v : SORT_NONE;
  }
#pragma unroll
  for (int k = 0; k < SI; ++k)
    rank[k] = atomicAdd(&st.cnt[sb[k] != SORT_NONE ? sb[k] >> SORT_COARSE_SHIFT
                                                  : (uint32_t)SORT_MAX_BUCKETS + (tid & (WAVE - 1))], 1u);
  lds_barrier();
  // 2. one lane per non-empty bucket: reserve the run, clear the count
  for (uint32_t b = tid; b < a.n_buckets; b += N) {
    const uint32_t c = st.cnt[b];
    if (c) {
      const uint32_t gb = atomicAdd(a.bcur + b, c);
      const uint32_t cap = st.bcap[b];
      st.run[b] = make_uint2(st.boff[b] + gb, cap > gb ? cap - gb : 0u);
      st.cnt[b] = 0u;
    }
  }
  lds_barrier();
  // 3. every record to its place in its bucket's run (the LDS reads of four
  //    samples issued together); records beyond the bucket's capacity go
  //    straight to the histogram.  No barrier follows: the next stage writes
  //    only this lane's own slots, and the next flush rewrites cnt and run
  //    only after its first barrier, which every wave reaches after these
  //    reads.
  constexpr int G = SI < 4 ? SI : 4;
  static_assert(SI % G == 0, "groups of samples");
  bool any_ovf = false;
#pragma unroll
  for (int k0 = 0; k0 < SI; k0 += G) {
    uint2 r[G];
    uint32_t p[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const uint32_t q = sb[k0 + k];
      r[k] = st.run[q != SORT_NONE ? q >> SORT_COARSE_SHIFT : 0u];
      p[k] = st.pid[(k0 + k) * N + tid];
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const uint32_t q = sb[k0 + k];
      const bool v = q != SORT_NONE;
      const bool fits = v && rank[k0 + k] < r[k].y;
      if (fits) a.cbuf[r[k].x + rank[k0 + k]] = ((q & ((1u << SORT_COARSE_SHIFT) - 1u)) << 8) | p[k];
      // a sample that does not fit keeps rank ~0 for the pass below
      rank[k0 + k] = v && !fits ? ~0u : 0u;
      any_ovf = any_ovf || (v && !fits);
    }
  }
  if (__builtin_amdgcn_ballot_w64(any_ovf)) {  // wave-uniform and rare
#pragma unroll
    for (int k = 0; k < SI; ++k)
      bin_quad(bins, sb[k], rank[k] == ~0u, st.pid[k * N + tid], 1.0f, pal);
  }
}

template <int WAVES, int HM, int LOGK, class Prog>
__device__ __forceinline__ void k_fast_body(const FastArgs& g, const Block* __restrict__ blocks) {
  constexpr int N = 64 * WAVES;
  // Double-buffered exchange: iteration `it` writes buffer it&1, so one
  // barrier per iteration separates every write from the reads it must not
  // overtake (the next write to the same buffer is two barriers later).
  __shared__ float s_x[2][N], s_y[2][N], s_c[2][N];
  // The sort path reads the palette only for the rare records beyond a
  // bucket's capacity, from global memory: its LDS stays within three
  // workgroups per CU.
  __shared__ float4 s_pal[HM == 5 ? 1 : 256];
  // The permutation rows live in LDS: a global load inside the loop would sit
  // behind the iteration's atomics in the in-order vmcnt queue and stall.
  __shared__ __attribute__((aligned(16))) uint16_t s_perm[NPERM * N];
  __shared__ LdsBins<LOGK> s_bins;
  __shared__ SortStage<N, HM == 5> s_sort;
  const uint32_t tid = threadIdx.x;
  if constexpr (HM != 5)
    for (uint32_t i = tid; i < 256u; i += N) s_pal[i] = g.palette[i];
  // 16-byte copies (the table is 16-byte aligned in device memory and in LDS)
  static_assert((NPERM * N) % 8 == 0, "whole 16-byte pieces");
  for (uint32_t i = tid; i < (uint32_t)(NPERM * N / 8); i += N)
    reinterpret_cast<uint4*>(s_perm)[i] = reinterpret_cast<const uint4*>(g.perms)[i];
  if constexpr (HM == 4) {
    lds_bins_clear(s_bins);
    lds_bins_palette(s_bins, g.palette);
  }
  if constexpr (HM == 5) {
    if (g.do_draw) {
      for (uint32_t b = tid; b < g.sort.n_buckets; b += N) {
        s_sort.boff[b] = g.sort.boff[b];
        s_sort.bcap[b] = g.sort.bcap[b];
        s_sort.cnt[b] = 0u;
      }
    }
  }
  __syncthreads();

  unsigned long long drawn = 0;
  uint32_t hits = 0;
  for (uint32_t batch = blockIdx.x; batch < g.n_batches; batch += gridDim.x) {
    const uint32_t base = batch * (uint32_t)N;
    const uint32_t pid = base + tid;
    // Integer division runs on the VALU, so the quotient is not known to be
    // uniform; without readfirstlane every Block access would be a vector load.
    const uint32_t t = __builtin_amdgcn_readfirstlane(base / g.ppt);
    CBlock* blk = as_constant(blocks + (size_t)t * g.block_stride);

    Jsf32 lrng = jsf_from_philox(g.seed, pid, g.launch, 0u);
    const uint32_t wid = __builtin_amdgcn_readfirstlane(batch * WAVES + (tid >> 6));
    Jsf32 wrng = jsf_from_philox(g.seed, wid, g.launch, 1u);
    Jsf32 grng = jsf_from_philox(g.seed, batch, g.launch, 2u);

    float x, y, c;
    if (g.first_run) {
      // Hammersley point (src/hammersley.cpp:29-48) of index pid mod ppt, jittered
      // like the first run of shaders/flame.glsl:63-66.
      const uint32_t cnt = g.ppt;
      const uint32_t mx = (cnt % 2u == 0u) ? cnt : (1u << (32 - __clz(cnt - 1u)));
      const float inv_max = rf_div_rn(1.0f, (float)mx);
      const uint32_t rbits = 31u - __clz(mx);
      const uint32_t i = pid % cnt;
      const uint32_t fl = __brev(i) >> (32u - rbits);
      x = ((float)i * inv_max) * 2.0f - 1.0f;
      y = ((float)fl * inv_max) * 2.0f - 1.0f;
      c = 0.0f;
      const float r1 = lrng.randf();
      const float r2 = lrng.randf();
      const float s = ((r1 * 0.1f) * PI) * 2.0f;
      float sn, cs;
      rf_sincos01(rf_sqrt_rn(r2), &sn, &cs);
      x = x + s * sn;
      y = y + s * cs;
    } else {
      const float4 p = g.pos[pid];
      x = p.x; y = p.y; c = p.z;
    }

    for (uint32_t it = 0; it < g.n_iter; ++it) {
      // The draw is wave-uniform; selecting on the raw integer keeps xid in SGPRs
      // (float compares would run on the VALU, make xid look divergent and turn
      // every xform/term access into vector loads behind the atomics).
      const int xid = Prog::select(blk, wrng.next_a());
      XOut o = Prog::apply(blk, xid, x, y, c, g.first_run && it == 0, lrng);
      if (rf_badval(o.x) || rf_badval(o.y)) {
        // A particle that left the finite plane restarts at a random point and
        // is not drawn this iteration (the reference leaves it NaN forever,
        // shaders/flame.glsl:70 is commented out).
        o.x = lrng.randf() * 2.0f - 1.0f;
        o.y = lrng.randf() * 2.0f - 1.0f;
        o.w = 0.0f;
      }
      x = o.x; y = o.y; c = o.c;
      if (g.do_draw) {
        const Located L = locate<Prog>(blk, o, lrng, g.ss, g.W, g.H);
        if constexpr (HM == 4) {
          bin_lds(s_bins, g.bins, L.bidx, L.valid, L.pidx, L.w, s_pal, hits);
        } else if constexpr (HM == 5) {
          const uint32_t slot = (it % sort_stage_iters<N>()) * N + tid;
          s_sort.bin[slot] = L.valid ? L.bidx : SORT_NONE;
          s_sort.pid[slot] = (uint8_t)L.pidx;
        } else {
          bin_sample<HM>(g.bins, L.bidx, L.valid, L.pidx, L.w, s_pal);
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
        if (g.do_draw && ((it % SI) == SI - 1 || it + 1 == g.n_iter))
          sort_flush(s_sort, (it % SI) + 1, g.sort, g.bins, g.palette);
      }
    }
    g.pos[pid] = make_float4(x, y, c, 0.0f);
    // The next batch's first exchange writes buffer 0, which slower waves may
    // still be reading from this batch's last iteration.
    lds_barrier();
  }
  if (g.do_draw) {
    count_drawn(g.counter, drawn);
    if constexpr (HM == 4) {
      __syncthreads();
      flush_lds(s_bins, g.bins);
      count_slot(g.counter, 1, (unsigned long long)wave_sum_u32(hits));
    }
  }
}

}  // namespace detail
}  // namespace rfk
