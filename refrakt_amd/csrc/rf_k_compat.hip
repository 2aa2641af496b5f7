// Compat engine: one launch per pass, the reference iterate shader restated for gfx950.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_kern_common.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

// ----------------------------------------------------------- compat pass
// One pass of the reference iterate shader (shaders/flame.glsl:41-90) over
// all P = ppt * T particles: grid (ppt/256, T), 256 lanes per workgroup, one
// xform drawn by lane 0 for the whole workgroup, per-slot jsf32 state,
// shuffle-row gather on read (and scatter on write during warmup).
template <int HM>
__global__ __launch_bounds__(256) void k_compat_pass(CompatArgs g, const Block* __restrict__ blocks) {
  __shared__ int s_xid;
  __shared__ float4 s_pal[256];
  const uint32_t gx = blockIdx.x * 256u + threadIdx.x;
  const uint32_t y = blockIdx.y;
  const size_t slot = (size_t)y * g.ppt + gx;
  const uint4 st = g.rs[slot];
  Jsf32 rng{st.x, st.y, st.z, st.w};
  CBlock* blk = as_constant(blocks + (size_t)y * g.block_stride);
  if (g.do_draw) s_pal[threadIdx.x] = g.palette[threadIdx.x];
  if (threadIdx.x == 0) s_xid = select_xform(blk, rng.randf());
  const uint32_t i_idx = g.random_read ? g.shuf[(size_t)g.in_row * g.ppt + gx] : gx;
  const uint32_t o_idx = g.random_write ? g.shuf[(size_t)g.out_row * g.ppt + gx] : gx;
  const size_t off = g.first_run ? 0 : (size_t)y * g.ppt;
  float4 p = g.pos_in[off + i_idx];
  if (g.first_run) {
    // part_state.xy += randf() * .1 * PI * 2.0 * sincos(sqrt(randf()))
    const float r1 = rng.randf();
    const float r2 = rng.randf();
    const float s = ((r1 * 0.1f) * PI) * 2.0f;
    float sn, cs;
    rf_sincos01(rf_sqrt_rn(r2), &sn, &cs);
    p.x = p.x + s * sn;
    p.y = p.y + s * cs;
  }
  __syncthreads();
  const int xid = __builtin_amdgcn_readfirstlane(s_xid);
  const XOut o = apply_xform(blk, xid, p.x, p.y, p.z, g.first_run != 0, rng);
  g.pos_out[(size_t)y * g.ppt + o_idx] = make_float4(o.x, o.y, o.c, 0.0f);
  if (g.do_draw) {
    const bool v = draw_point<HM>(blk, o, rng, g.ss, g.W, g.H, g.bins, s_pal);
    count_drawn(g.counter, (unsigned long long)__popcll(__ballot(v)));
  }
  g.rs[slot] = make_uint4(rng.a, rng.b, rng.c, rng.d);
}

// jsf32::warmup_ctx(state[i], i) for every slot (src/flame.cpp:136-141).
__global__ void k_init_rs(uint4* rs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Jsf32 s = Jsf32::seeded((uint32_t)i);
  rs[i] = make_uint4(s.a, s.b, s.c, s.d);
}

}  // namespace

hipError_t launch_init_rs(uint4* rs, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_init_rs, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, rs, n);
  return hipGetLastError();
}

hipError_t launch_compat_pass(const CompatArgs& a, int hist_mode, hipStream_t s) {
  dim3 grid(a.ppt / 256u, a.T), block(256);
  switch (hist_mode) {
    case 2: hipLaunchKernelGGL(k_compat_pass<2>, grid, block, 0, s, a, a.blocks); break;
    case 3: hipLaunchKernelGGL(k_compat_pass<3>, grid, block, 0, s, a, a.blocks); break;
    default: hipLaunchKernelGGL(k_compat_pass<1>, grid, block, 0, s, a, a.blocks); break;
  }
  return hipGetLastError();
}

}  // namespace rfk
