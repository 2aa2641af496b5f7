// Binning and drawing helpers shared by the chaos-game kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

template <int HM>
__device__ __forceinline__ void bin_sample(float* __restrict__ bins, uint32_t bidx, bool valid,
                                           uint32_t pidx, float w, const float4* __restrict__ pal) {
  if constexpr (HM == 2) bin_waveagg(bins, bidx, valid, pidx, w, pal);
  else if constexpr (HM == 3) bin_naive(bins, bidx, valid, pidx, w, pal);
  else bin_quad(bins, bidx, valid, pidx, w, pal);
}

// Final xform, screen affine, floor, bounds and opacity test, then the bin
// add (shaders/flame.glsl:74-87).  Float comparisons replace the integer
// ones of the reference: floor(s) >= 0 <=> s >= 0 and floor(s) < W <=> s < W,
// and NaN fails both (the reference's int conversion of NaN is undefined).
template <int HM, class Rng, class BP>
__device__ __forceinline__ bool draw_point(BP blk, const XOut& o, Rng& rng,
                                           const float ss[6], uint32_t W, uint32_t H,
                                           float* __restrict__ bins, const float4* __restrict__ pal) {
  float rx = o.x, ry = o.y, rc = o.c, rw = o.w;
  if (blk->has_final) {
    const XOut f = apply_xform(blk, FINAL, o.x, o.y, o.c, false, rng);
    rx = f.x; ry = f.y; rc = f.c; rw = f.w * o.w;
  }
  const float sx = fmaf(ss[0], rx, fmaf(ss[2], ry, ss[4]));
  const float sy = fmaf(ss[1], rx, fmaf(ss[3], ry, ss[5]));
  const bool valid = (sx >= 0.0f) && (sx < (float)W) && (sy >= 0.0f) && (sy < (float)H) && (rw > 0.0f);
  uint32_t bidx = 0;
  if (valid) {
    const uint32_t cx = (uint32_t)floorf(sx), cy = (uint32_t)floorf(sy);
    bidx = (H - cy - 1u) * W + cx;
  }
  bin_sample<HM>(bins, bidx, valid, rf_pal_index(rc), rw, pal);
  return valid;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global access (vmcnt(0)), which would stall each
// iteration on the round trip of its fire-and-forget histogram atomics.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void count_drawn(unsigned long long* counter, unsigned long long n) {
  const int lane = threadIdx.x & (WAVE - 1);
  if (lane == 0 && n) atomicAdd(counter, n);
}

}  // namespace detail
}  // namespace rfk
