// Binning and drawing helpers shared by the chaos-game kernels.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "rf_internal.hpp"

namespace rfk {
namespace detail {
using namespace rfd;

constexpr int WAVE = 64;

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
  // Relaxed, agent scope, result unused: one global_atomic_add_f32 without return.
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------- binning
// Each wave holds up to 64 samples (one per lane).  A sample adds
// (palette[pidx].rgb, w) to the 16-byte bin `bidx`.
//
// HIST_QUAD: the 4 channels of one sample are added by 4 neighbouring lanes
// in the same instruction, so the 4 dwords travel as one 16-byte atomic
// request to L2 instead of 4 requests.  On MI355X the L2 atomic path is
// bound by requests (about 21 G/s measured, up to 16 dwords each), so this
// is 4x the naive rate for random bins.
__device__ __forceinline__ void bin_quad(float* __restrict__ bins, uint32_t bidx, bool valid,
                                         uint32_t pidx, float w, const float4* __restrict__ pal) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t vm = __ballot(valid);
  if (vm == 0) return;
  const float* palf = reinterpret_cast<const float*>(pal);
  const int ch = lane & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t grp = (uint32_t)(vm >> (16 * k)) & 0xffffu;
    if (grp == 0) continue;  // wave-uniform
    const int src = 16 * k + (lane >> 2);
    const uint32_t b = (uint32_t)__shfl((int)bidx, src);
    const uint32_t p = (uint32_t)__shfl((int)pidx, src);
    const float ww = __shfl(w, src);
    if ((grp >> (lane >> 2)) & 1u) {
      const float v = (ch == 3) ? ww : palf[p * 4 + ch];
      atomic_add_f32(bins + (size_t)b * 4 + ch, v);
    }
  }
}

// Baseline: each lane issues the 4 channel atomics of its own sample.
__device__ __forceinline__ void bin_naive(float* __restrict__ bins, uint32_t bidx, bool valid,
                                          uint32_t pidx, float w, const float4* __restrict__ pal) {
  if (!valid) return;
  const float4 c = pal[pidx];
  float* p = bins + (size_t)bidx * 4;
  atomic_add_f32(p + 0, c.x);
  atomic_add_f32(p + 1, c.y);
  atomic_add_f32(p + 2, c.z);
  atomic_add_f32(p + 3, w);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

// Exclusive prefix sum of cnt[0..n) into off[0..n) (n <= 64 * PER), and the
// total into *total; executed by the first wave of the workgroup only.
template <int PER>
__device__ __forceinline__ void wave0_exclusive_scan(const uint32_t* cnt, uint32_t* off, uint32_t n,
                                                     uint32_t* total) {
  if (threadIdx.x >= (uint32_t)WAVE) return;
  const uint32_t lane = threadIdx.x;
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = lane * PER + k;
    v[k] = i < n ? cnt[i] : 0u;
    s += v[k];
  }
  uint32_t inc = s;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)inc, d);
    if ((int)lane >= d) inc += t;
  }
  uint32_t run = inc - s;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = lane * PER + k;
    if (i < n) off[i] = run;
    run += v[k];
  }
  if (lane == WAVE - 1) *total = inc;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// HIST_WAVEAGG: lanes that hit the same bin are merged first (one leader per
// distinct bin adds the wave's sum), which removes same-address serialisation
// at L2 when a flame concentrates its hits.  While the leaders keep finding
// singleton bins the remaining lanes fall through to the quad path.
__device__ __forceinline__ void bin_waveagg(float* __restrict__ bins, uint32_t bidx, bool valid,
                                            uint32_t pidx, float w, const float4* __restrict__ pal) {
  const int lane = threadIdx.x & (WAVE - 1);
  uint64_t rem = __ballot(valid);
  if (rem == 0) return;
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) { const float4 q = pal[pidx]; c = make_float4(q.x, q.y, q.z, w); }
  int singletons = 0;
  while (rem) {
    const int leader = __ffsll((unsigned long long)rem) - 1;
    const uint32_t key = (uint32_t)__shfl((int)bidx, leader);
    const bool mine = ((rem >> lane) & 1ull) && bidx == key;
    const uint64_t m = __ballot(mine);
    if (__popcll(m) == 1) {
      if (++singletons >= 2) break;  // sparse footprint: quad path is cheaper
    }
    const float sx = wave_sum(mine ? c.x : 0.f);
    const float sy = wave_sum(mine ? c.y : 0.f);
    const float sz = wave_sum(mine ? c.z : 0.f);
    const float sw = wave_sum(mine ? c.w : 0.f);
    if (lane < 4) {
      const float v = lane == 0 ? sx : lane == 1 ? sy : lane == 2 ? sz : sw;
      atomic_add_f32(bins + (size_t)key * 4 + lane, v);
    }
    rem &= ~m;
  }
  bin_quad(bins, bidx, (rem >> lane) & 1ull, pidx, w, pal);
}

// Adds one RGBA value per lane (`v`) to bin `bidx` with the quad layout of
// bin_quad: 4 lanes per bin, one 16-byte request per value.
__device__ __forceinline__ void bin_quad_vals(float* __restrict__ bins, uint32_t bidx, bool valid, float4 v) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t vm = __ballot(valid);
  if (vm == 0) return;
  const int ch = lane & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t grp = (uint32_t)(vm >> (16 * k)) & 0xffffu;
    if (grp == 0) continue;
    const int src = 16 * k + (lane >> 2);
    const uint32_t b = (uint32_t)__shfl((int)bidx, src);
    const float vx = __shfl(v.x, src), vy = __shfl(v.y, src), vz = __shfl(v.z, src), vw = __shfl(v.w, src);
    if ((grp >> (lane >> 2)) & 1u) {
      const float val = ch == 0 ? vx : ch == 1 ? vy : ch == 2 ? vz : vw;
      atomic_add_f32(bins + (size_t)b * 4 + ch, val);
    }
  }
}

// Fixed-point LDS accumulation.  LDS float atomics are executed about 25x
// slower than integer ones on gfx950 (measured: ds_add_f32 0.2 T, ds_add_u32
// 5.0 T, ds_add_u64 3.0 T lane-ops/s), so LDS sums are kept as u64 multiples
// of 2^-31: exact for values in [0, 1] quantised to 2^-31, and 2^33 samples
// of 1.0 fit.  Palette entries and weights outside [0, 1] are not accepted
// by the modes that use this (the host falls back to the quad path).
constexpr double FIX_ONE = 2147483648.0;  // 2^31
constexpr double FIX_INV = 1.0 / 2147483648.0;

__device__ __forceinline__ uint64_t to_fix(float v) { return (uint64_t)((double)v * FIX_ONE); }
__device__ __forceinline__ float from_fix(uint64_t q) { return (float)((double)q * FIX_INV); }

__device__ __forceinline__ void lds_add_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// HIST_LDS: each workgroup keeps an open-addressed table of 2^LOGK bins in
// LDS (key = bin + 1, 0 = free; claimed on first touch, never evicted) and
// adds the samples that find or claim a slot there with LDS integer atomics
// (fixed point, above); the rest take the quad path.  The table is added to
// the histogram once, when the workgroup finishes (flush_lds), so a bin hit
// thousands of times by one workgroup costs one global request instead of
// thousands.  This is what a concentrated footprint needs: same-address
// atomics serialise at the memory-side atomic unit (a single hot bin sustains
// under 0.1 G adds/s).
constexpr int LDS_PROBES = 4;

// Values are four channel planes (not one record per slot) so that a wave's
// atomics to random slots spread over the LDS banks.
template <int LOGK>
struct LdsBins {
  uint32_t key[1u << LOGK];
  uint64_t val[4][1u << LOGK];
  uint32_t palq[3][256];  // palette in fixed point
};
template <>
struct LdsBins<0> {};  // modes without the cache: no LDS

template <int LOGK>
__device__ __forceinline__ void lds_bins_clear(LdsBins<LOGK>& t) {
  for (uint32_t i = threadIdx.x; i < (1u << LOGK); i += blockDim.x) {
    t.key[i] = 0u;
    t.val[0][i] = 0ull;
    t.val[1][i] = 0ull;
    t.val[2][i] = 0ull;
    t.val[3][i] = 0ull;
  }
}

// Fills the fixed-point palette from a 256-entry float palette in global memory.
template <int LOGK>
__device__ __forceinline__ void lds_bins_palette(LdsBins<LOGK>& t, const float4* __restrict__ pal) {
  for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) {
    const float4 c = pal[i];
    t.palq[0][i] = (uint32_t)to_fix(c.x);
    t.palq[1][i] = (uint32_t)to_fix(c.y);
    t.palq[2][i] = (uint32_t)to_fix(c.z);
  }
}

// `hits` counts the samples whose bin was already in the table (the claim
// of a free slot is not counted): the measure of what the table saves.
template <int LOGK>
__device__ __forceinline__ void bin_lds(LdsBins<LOGK>& t, float* __restrict__ bins, uint32_t bidx, bool valid,
                                        uint32_t pidx, float w, const float4* __restrict__ pal, uint32_t& hits) {
  constexpr uint32_t K = 1u << LOGK;
  int slot = -1;
  bool claimed = false;
  if (valid) {
    const uint32_t key = bidx + 1u;
    uint32_t s = (bidx * 0x9E3779B1u) >> (32 - LOGK);
    for (int p = 0; p < LDS_PROBES; ++p) {
      uint32_t k = __hip_atomic_load(&t.key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (k == 0u) {
        uint32_t expected = 0u;
        claimed = __hip_atomic_compare_exchange_strong(&t.key[s], &expected, key, __ATOMIC_RELAXED,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        k = claimed ? key : expected;
      }
      if (k == key) { slot = (int)s; break; }
      s = (s + 1u) & (K - 1u);
    }
  }
  if (slot >= 0) {
    lds_add_u64(&t.val[0][slot], (uint64_t)t.palq[0][pidx]);
    lds_add_u64(&t.val[1][slot], (uint64_t)t.palq[1][pidx]);
    lds_add_u64(&t.val[2][slot], (uint64_t)t.palq[2][pidx]);
    lds_add_u64(&t.val[3][slot], w == 1.0f ? (uint64_t)FIX_ONE : to_fix(w));
    // Only re-hits count: a table that is merely being filled saves nothing.
    if (!claimed) ++hits;
  }
  bin_quad(bins, bidx, valid && slot < 0, pidx, w, pal);
}

// Adds every claimed slot to the histogram.  Call after a full barrier; the
// loop bound is workgroup-uniform so every wave takes part in every ballot.
template <int LOGK>
__device__ __forceinline__ void flush_lds(LdsBins<LOGK>& t, float* __restrict__ bins) {
  for (uint32_t s0 = 0; s0 < (1u << LOGK); s0 += blockDim.x) {
    const uint32_t s = s0 + threadIdx.x;
    const uint32_t k = s < (1u << LOGK) ? t.key[s] : 0u;
    const float4 v = k ? make_float4(from_fix(t.val[0][s]), from_fix(t.val[1][s]), from_fix(t.val[2][s]),
                                     from_fix(t.val[3][s]))
                       : make_float4(0.f, 0.f, 0.f, 0.f);
    bin_quad_vals(bins, k - 1u, k != 0u, v);
  }
}

template <int HM>
__device__ __forceinline__ void bin_sample(float* __restrict__ bins, uint32_t bidx, bool valid,
                                           uint32_t pidx, float w, const float4* __restrict__ pal) {
  if constexpr (HM == 2) bin_waveagg(bins, bidx, valid, pidx, w, pal);
  else if constexpr (HM == 3) bin_naive(bins, bidx, valid, pidx, w, pal);
  else bin_quad(bins, bidx, valid, pidx, w, pal);
}

// Final xform, screen affine, floor, bounds and opacity test
// (shaders/flame.glsl:74-87).  Float comparisons replace the integer ones of
// the reference: floor(s) >= 0 <=> s >= 0 and floor(s) < W <=> s < W, and NaN
// fails both (the reference's int conversion of NaN is undefined).
struct Located {
  uint32_t bidx;   // (H - 1 - floor(sy)) * W + floor(sx); 0 when !valid
  uint32_t pidx;   // palette index
  float w;         // opacity weight
  bool valid;
};

// How a chaos-game kernel applies a flame's xforms.  InterpProg reads the
// structure (variations, precalcs, post-affine, final xform) from the Block at
// run time; a compiled flame program (rf_jit.cpp) supplies the same four
// functions with the structure fixed at compile time.
struct InterpProg {
  template <class BP>
  static __device__ __forceinline__ int select(BP blk, uint32_t raw) { return select_xform_raw(blk, raw); }
  template <class Rng, class BP>
  static __device__ __forceinline__ XOut apply(BP blk, int xi, float x, float y, float c, bool first, Rng& rng) {
    return apply_xform(blk, xi, x, y, c, first, rng);
  }
  template <class BP>
  static __device__ __forceinline__ bool has_final(BP blk) { return blk->has_final != 0u; }
  template <class Rng, class BP>
  static __device__ __forceinline__ XOut apply_final(BP blk, float x, float y, float c, Rng& rng) {
    return apply_xform(blk, FINAL, x, y, c, false, rng);
  }
};

template <class Prog = InterpProg, class Rng, class BP>
__device__ __forceinline__ Located locate(BP blk, const XOut& o, Rng& rng, const float ss[6], uint32_t W,
                                          uint32_t H) {
  float rx = o.x, ry = o.y, rc = o.c, rw = o.w;
  if (Prog::has_final(blk)) {
    const XOut f = Prog::apply_final(blk, o.x, o.y, o.c, rng);
    rx = f.x; ry = f.y; rc = f.c; rw = f.w * o.w;
  }
  const float sx = fmaf(ss[0], rx, fmaf(ss[2], ry, ss[4]));
  const float sy = fmaf(ss[1], rx, fmaf(ss[3], ry, ss[5]));
  Located L;
  // straight-line code: the bin is computed for every lane and kept only
  // where valid (a conversion of an out-of-range value is harmless there)
  L.valid = (sx >= 0.0f) & (sx < (float)W) & (sy >= 0.0f) & (sy < (float)H) & (rw > 0.0f);
  const uint32_t cx = (uint32_t)floorf(sx), cy = (uint32_t)floorf(sy);
  L.bidx = L.valid ? (H - cy - 1u) * W + cx : 0u;
  L.pidx = rf_pal_index(rc);
  L.w = rw;
  return L;
}

template <int HM, class Rng, class BP>
__device__ __forceinline__ bool draw_point(BP blk, const XOut& o, Rng& rng,
                                           const float ss[6], uint32_t W, uint32_t H,
                                           float* __restrict__ bins, const float4* __restrict__ pal) {
  const Located L = locate(blk, o, rng, ss, W, H);
  bin_sample<HM>(bins, L.bidx, L.valid, L.pidx, L.w, pal);
  return L.valid;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global access (vmcnt(0)), which would stall each
// iteration on the round trip of its fire-and-forget histogram atomics.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Adds a wave-uniform count to field `field` of this wave's counter slot.
__device__ __forceinline__ void count_slot(unsigned long long* counter, int field, unsigned long long n) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const uint32_t slot = (blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) % (uint32_t)COUNTER_SLOTS;
  if (lane == 0 && n) atomicAdd(counter + slot * COUNTER_STRIDE + field, n);
}
__device__ __forceinline__ void count_drawn(unsigned long long* counter, unsigned long long n) {
  count_slot(counter, 0, n);
}

}  // namespace detail
}  // namespace rfk
