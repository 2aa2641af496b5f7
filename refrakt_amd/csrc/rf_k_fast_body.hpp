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
// flush ranks them by bucket (LDS counting sort; each lane loads its staged
// bins first, then issues its rank atomics back to back and keeps the ranks
// in registers), wave 0 scans the counts, reserves one run per non-empty
// bucket with one global atomic and clears the counts, the records are placed
// in bucket order in LDS and written out so that neighbouring lanes store
// neighbouring addresses.  Four barriers per flush.
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
  // Each lane reads back only its own staged slots, into registers, at the
  // start of a flush; the bucket-ordered records of the flush then reuse the
  // memory (the flush ends with a barrier before the next stage writes).
  // rec and bkt have a spare entry S: the placement of an empty slot writes
  // there, and the write-out reads it for positions past the last record, so
  // neither needs a branch around its LDS accesses.
  union {
    uint32_t bin[S];     // staged bin index (SORT_NONE: no sample), slot = iteration * N + lane
    uint32_t rec[S + 1]; // bucket-ordered records
  };
  uint8_t pid[S];        // staged palette index
  uint8_t bkt[S + 1];    // bucket of rec[j]
  uint32_t cnt[SORT_MAX_BUCKETS];    // records per bucket in this flush (zero between flushes)
  uint32_t off[SORT_MAX_BUCKETS];    // exclusive scan of cnt
  uint32_t base[SORT_MAX_BUCKETS];   // record j of bucket b goes to cbuf[base[b] + j] ...
  uint32_t lim[SORT_MAX_BUCKETS];    // ... if j < lim[b], else over the bucket's capacity
  uint32_t boff[SORT_MAX_BUCKETS];   // bucket offsets and capacities, staged once per workgroup
  uint32_t bcap[SORT_MAX_BUCKETS];
  uint32_t total;
};
template <int N>
struct SortStage<N, false> {};

template <int N>
__device__ __forceinline__ void sort_flush(SortStage<N, true>& st, uint32_t n_staged, const SortArgs& a,
                                           float* __restrict__ bins, const float4* __restrict__ pal) {
  constexpr int SI = sort_stage_iters<N>();
  static_assert(SORT_MAX_BUCKETS % WAVE == 0, "scan segments");
  const uint32_t tid = threadIdx.x;
  // 1. rank this lane's staged samples within their buckets
  uint32_t sb[SI], rank[SI];
#pragma unroll
  for (int k = 0; k < SI; ++k) sb[k] = (uint32_t)k < n_staged ? st.bin[k * N + tid] : SORT_NONE;
#pragma unroll
  for (int k = 0; k < SI; ++k)
    rank[k] = sb[k] != SORT_NONE ? atomicAdd(&st.cnt[sb[k] >> SORT_COARSE_SHIFT], 1u) : 0u;
  lds_barrier();
  // 2. wave 0: exclusive scan of the counts (64-bucket segments, lane = bucket
  //    within the segment), one reservation atomic per non-empty bucket,
  //    destinations, and the counts cleared for the next flush
  if (tid < (uint32_t)WAVE) {
    constexpr int SEG = SORT_MAX_BUCKETS / WAVE;
    uint32_t v[SEG], gb[SEG];
    // all reservation atomics first (one round trip), then the scans
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
      const uint32_t b = k * WAVE + tid;
      v[k] = b < a.n_buckets ? st.cnt[b] : 0u;
      gb[k] = v[k] ? atomicAdd(a.bcur + b, v[k]) : 0u;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
      if ((uint32_t)(k * WAVE) >= a.n_buckets) break;  // uniform
      const uint32_t b = k * WAVE + tid;
      uint32_t inc = v[k];
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)inc, d);
        if ((int)tid >= d) inc += t;
      }
      const uint32_t run = carry + inc - v[k];
      if (b < a.n_buckets) {
        const uint32_t cap = st.bcap[b];
        st.off[b] = run;
        st.base[b] = st.boff[b] + gb[k] - run;
        st.lim[b] = run + (cap > gb[k] ? cap - gb[k] : 0u);
        st.cnt[b] = 0u;
      }
      carry += (uint32_t)__shfl((int)inc, WAVE - 1);
    }
    if (tid == 0) st.total = carry;
  }
  lds_barrier();
  // 3. place the records in bucket order (all reads first, then the writes;
  //    an empty slot goes to the spare entry)
  constexpr uint32_t S = (uint32_t)SortStage<N, true>::S;
  uint32_t pb[SI], po[SI];
#pragma unroll
  for (int k = 0; k < SI; ++k) {
    pb[k] = sb[k] != SORT_NONE ? sb[k] >> SORT_COARSE_SHIFT : 0u;
    po[k] = st.off[pb[k]];
    rank[k] |= (uint32_t)st.pid[k * N + tid] << 24;  // ranks stay below 2^13
  }
#pragma unroll
  for (int k = 0; k < SI; ++k) {
    const uint32_t j = sb[k] != SORT_NONE ? po[k] + (rank[k] & 0xffffffu) : S;
    st.rec[j] = ((sb[k] & ((1u << SORT_COARSE_SHIFT) - 1u)) << 8) | (rank[k] >> 24);
    st.bkt[j] = (uint8_t)pb[k];
  }
  lds_barrier();
  // 4. coalesced write-out; records beyond a bucket's capacity go straight to
  //    the histogram.  The barrier at the end keeps the next stage's writes
  //    (to the slots rec shares) behind every wave's reads; bkt, base, lim
  //    and total are next written after the next flush's first barrier.
  //    Two rounds at a time, their LDS reads issued together; positions past
  //    the last record read the spare entry and store nothing.
  const uint32_t total = st.total;
  const uint32_t rounds = (total + N - 1) / N;  // uniform
  for (uint32_t r = 0; r < rounds; r += 2) {
    uint32_t j[2], b[2], rc[2], lim[2], base[2];
    bool in[2], ovf[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      j[h] = (r + (uint32_t)h) * N + tid;
      in[h] = j[h] < total;
      const uint32_t jc = in[h] ? j[h] : S;
      b[h] = st.bkt[jc];
      rc[h] = st.rec[jc];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      lim[h] = st.lim[b[h]];
      base[h] = st.base[b[h]];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      ovf[h] = in[h] && j[h] >= lim[h];
      if (in[h] && !ovf[h]) a.cbuf[base[h] + j[h]] = rc[h];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
      bin_quad(bins, (b[h] << SORT_COARSE_SHIFT) | (rc[h] >> 8), ovf[h], rc[h] & 255u, 1.0f, pal);
  }
  lds_barrier();
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
  __shared__ uint16_t s_perm[NPERM * N];
  __shared__ LdsBins<LOGK> s_bins;
  __shared__ SortStage<N, HM == 5> s_sort;
  const uint32_t tid = threadIdx.x;
  if constexpr (HM != 5)
    for (uint32_t i = tid; i < 256u; i += N) s_pal[i] = g.palette[i];
  for (uint32_t i = tid; i < (uint32_t)(NPERM * N); i += N) s_perm[i] = g.perms[i];
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
