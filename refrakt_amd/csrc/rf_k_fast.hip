// Fast engine: persistent chaos game with wave-uniform xforms and an LDS particle exchange.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_kern_common.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

// ------------------------------------------------------------ fast engine
// Persistent chaos game: a workgroup owns 64*WAVES particles for all n_iter
// iterations.  Per iteration each wave draws one xform from its own
// wave-uniform stream (so the variation switch never diverges), applies it,
// bins, and then all particles of the workgroup are permuted through LDS with
// one of NPERM random permutations chosen by a workgroup-uniform stream,
// which decorrelates the xform sequences of particles sharing a wave (the
// role of the reference's shuffle rows, src/shuffle_buffers.cpp).
template <int WAVES, int HM>
__global__ __launch_bounds__(64 * WAVES) void k_fast(FastArgs g, const Block* __restrict__ blocks) {
  constexpr int N = 64 * WAVES;
  // Double-buffered exchange: iteration `it` writes buffer it&1, so one
  // barrier per iteration separates every write from the reads it must not
  // overtake (the next write to the same buffer is two barriers later).
  __shared__ float s_x[2][N], s_y[2][N], s_c[2][N];
  __shared__ float4 s_pal[256];
  // The permutation rows live in LDS: a global load inside the loop would sit
  // behind the iteration's atomics in the in-order vmcnt queue and stall.
  __shared__ uint16_t s_perm[NPERM * N];
  const uint32_t tid = threadIdx.x;
  const uint32_t base = blockIdx.x * (uint32_t)N;
  const uint32_t pid = base + tid;
  // Integer division runs on the VALU, so the quotient is not known to be
  // uniform; without readfirstlane every Block access would be a vector load.
  const uint32_t t = __builtin_amdgcn_readfirstlane(base / g.ppt);
  CBlock* blk = as_constant(blocks + (size_t)t * g.block_stride);
  for (uint32_t i = tid; i < 256u; i += N) s_pal[i] = g.palette[i];
  for (uint32_t i = tid; i < (uint32_t)(NPERM * N); i += N) s_perm[i] = g.perms[i];

  Jsf32 lrng = jsf_from_philox(g.seed, pid, g.launch, 0u);
  const uint32_t wid = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + (tid >> 6));
  Jsf32 wrng = jsf_from_philox(g.seed, wid, g.launch, 1u);
  Jsf32 grng = jsf_from_philox(g.seed, blockIdx.x, g.launch, 2u);

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
  __syncthreads();

  unsigned long long drawn = 0;
  for (uint32_t it = 0; it < g.n_iter; ++it) {
    // The draw is wave-uniform; selecting on the raw integer keeps xid in SGPRs
    // (float compares would run on the VALU, make xid look divergent and turn
    // every xform/term access into vector loads behind the atomics).
    const int xid = select_xform_raw(blk, wrng.next_a());
    XOut o = apply_xform(blk, xid, x, y, c, g.first_run && it == 0, lrng);
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
      const bool v = draw_point<HM>(blk, o, lrng, g.ss, g.W, g.H, g.bins, s_pal);
      drawn += (unsigned long long)__popcll(__ballot(v));
    }
    const uint32_t row = grng.next_a() & (uint32_t)(NPERM - 1);
    const uint32_t dst = s_perm[row * N + tid];
    const uint32_t b = it & 1u;
    s_x[b][dst] = x; s_y[b][dst] = y; s_c[b][dst] = c;
    lds_barrier();
    x = s_x[b][tid]; y = s_y[b][tid]; c = s_c[b][tid];
  }
  g.pos[pid] = make_float4(x, y, c, 0.0f);
  if (g.do_draw) count_drawn(g.counter, drawn);
}

}  // namespace

template <int W>
static void launch_fast_w(const FastArgs& a, uint64_t n, int hm, hipStream_t s) {
  dim3 grid((uint32_t)(n / (64u * W))), block(64 * W);
  switch (hm) {
    case 2: hipLaunchKernelGGL((k_fast<W, 2>), grid, block, 0, s, a, a.blocks); break;
    case 3: hipLaunchKernelGGL((k_fast<W, 3>), grid, block, 0, s, a, a.blocks); break;
    default: hipLaunchKernelGGL((k_fast<W, 1>), grid, block, 0, s, a, a.blocks); break;
  }
}

hipError_t launch_fast(const FastArgs& a, uint64_t n_particles, int waves, int hist_mode, hipStream_t s) {
  switch (waves) {
    case 8: launch_fast_w<8>(a, n_particles, hist_mode, s); break;
    default: launch_fast_w<4>(a, n_particles, hist_mode, s); break;
  }
  return hipGetLastError();
}

}  // namespace rfk
