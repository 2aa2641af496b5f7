// Parameter inflation, density estimation + tonemap, diagnostics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_kern_common.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

// ---------------------------------------------------------------- animate
// Temporal-sample parameter inflation (shaders/templates/animate.tpl.glsl:18-96):
// block i (one workgroup each) is a copy of the pristine block blocks[T]
// with every xform's affine rotated by theta = 0.31415926535 * dt * rot[k],
// dt = (i - T/2) / float(T/2) * tss:
//   a' = a cos + c sin, b' = b cos + d sin, c' = c cos - a sin, d' = d cos - b sin.
// rot[k] is xform k's rotation_frequency (rot[n_xforms] for the final xform).
// The reference dispatches T/32 groups of 32 lanes, so it defines T >= 32
// only; at T = 1 (half = 0) the single block is the centre sample, dt = 0.
// The division is correctly rounded and sin/cos are evaluated in double and
// rounded to float, so the blocks are a deterministic function of the inputs
// that a CPU computes to the same bits.
__global__ __launch_bounds__(256) void k_animate(Block* blocks, uint32_t T, float tss,
                                                 const float* __restrict__ rot) {
  const uint32_t i = blockIdx.x;
  const Block* src = blocks + T;
  Block* dst = blocks + i;
  const uint32_t* s32 = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d32 = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t k = threadIdx.x; k < sizeof(Block) / 4; k += blockDim.x) d32[k] = s32[k];
  __syncthreads();
  const int nx = (int)src->n_xforms;
  const int k = (int)threadIdx.x;
  if (!(k < nx || (k == nx && src->has_final))) return;
  const int slot = k < nx ? k : FINAL;
  const int half = (int)(T / 2u);
  const int pos = (int)i - half;
  const float dt = half ? rf_div_rn((float)pos, (float)half) * tss : 0.0f;
  const float rps = 0.31415926535f;
  const float th = rps * dt * rot[k];
  const float sn = (float)sin((double)th), cs = (float)cos((double)th);
  const float* a = src->xf[slot].aff;
  float* b = dst->xf[slot].aff;
  const float t0 = a[0] * cs, t1 = a[2] * sn;
  const float t2 = a[1] * cs, t3 = a[3] * sn;
  const float t4 = a[2] * cs, t5 = a[0] * sn;
  const float t6 = a[3] * cs, t7 = a[1] * sn;
  b[0] = t0 + t1;
  b[1] = t2 + t3;
  b[2] = t4 - t5;
  b[3] = t6 - t7;
}

// ------------------------------------------------- density estimation
// Gather form of the reference's point-sprite splat (shaders/density_vert.glsl,
// shaders/density_frag.glsl, src/main.cpp:490-515).  A bin (px, py) with
// alpha a > 0 becomes a sprite centred on render-target texel
// (cx, cy) = (px - 1, H - 1 - py) (glm::ortho(0,W,H,0) maps its vertex at
// (px - 1/(4r+2), py + 1/(4r+2)) there; bins of column 0 have their centre
// outside the clip volume and are dropped).  Texel (cx + j, cy + i) receives
// color * (1 - d) * (2/pi)/r^2 where d = ((2j + 1/S)^2 + (2i + 1/S)^2) / S^2,
// S = 2r + 1, when |i|,|j| <= r and d <= 1; r = 0 copies the bin.
// r = max(rmin, min(R, int(R / a^curve))).
//
// Each workgroup computes a TW x TH tile of texels from an LDS copy of the
// (TW + 2R) x (TH + 2R) sources around it, then applies the tonemap.
constexpr int DE_TW = 32, DE_TH = 8;

struct DESrc { float r, e, nrm, pad; };

__device__ __forceinline__ void de_source(const DEArgs& g, int cx, int cy, float4& col, DESrc& m) {
  col = make_float4(0.f, 0.f, 0.f, 0.f);
  m.r = -1.0f; m.e = 0.0f; m.nrm = 0.0f; m.pad = 0.0f;
  const int px = cx + 1, py = (int)g.H - 1 - cy;
  if (cx < 0 || px >= (int)g.W || cy < 0 || cy >= (int)g.H) return;
  const float4 b = g.hist[(size_t)py * g.W + px];
  if (!(b.w != 0.0f)) return;  // color.a == 0.0 -> point moved out of the view volume
  const float q = rf_div_rn((float)g.R, powf(b.w, g.curve));
  int ri = (q >= (float)g.R) ? g.R : (int)q;
  if (!(q == q)) ri = g.R;
  int r = ri > g.rmin ? ri : g.rmin;
  if (r < 0) r = 0;
  col = b;
  m.r = (float)r;
  const int S = 2 * r + 1;
  m.e = rf_div_rn(1.0f, (float)S);
  m.nrm = (r > 0) ? rf_div_rn(0.63661977236f, (float)(r * r)) : 1.0f;
}

__device__ __forceinline__ float de_weight(const DESrc& m, int i, int j) {
  const int r = (int)m.r;
  if (r < 0) return 0.0f;
  if (r == 0) return (i == 0 && j == 0) ? 1.0f : 0.0f;
  if (i > r || i < -r || j > r || j < -r) return 0.0f;
  const float nx = (2.0f * (float)j + m.e) * m.e;
  const float ny = (2.0f * (float)i + m.e) * m.e;
  const float d = nx * nx + ny * ny;
  if (d > 1.0f) return 0.0f;
  return (1.0f - d) * m.nrm;
}

// shaders/tonemap.glsl:18-39 with scale_constant k; an empty texel (a = 0,
// where the shader divides 0/0) maps to black.
__device__ __forceinline__ float4 tonemap_px(float4 c, float gamma, float brightness, float vib, float k) {
  if (!(c.w > 0.0f)) return make_float4(0.f, 0.f, 0.f, 1.0f);
  // log(1 + a k) through log1pf: the shader's float log(1.0 + a*k) loses most
  // digits when a*k is small (k = 1e-4 makes that every sparse texel).
  const float ls = rf_div_rn(((0.5f * brightness) * log1pf(c.w * k)) * LOG10E, c.w);
  c.x *= ls; c.y *= ls; c.z *= ls; c.w *= ls;
  const float ig = rf_div_rn(1.0f, gamma);
  const float z = powf(c.w, ig);
  const float gf = rf_div_rn(z, c.w);
  float o[3] = {c.x, c.y, c.z};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float v = rf_mix(powf(o[q], ig), gf * o[q], vib);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    if (!(v == v)) v = 0.0f;
    o[q] = v;
  }
  return make_float4(o[0], o[1], o[2], 1.0f);
}

// Batched rendering: the last workgroup of a frame's DE stamps its completion.
// Every thread reaches this point (no early return), so the barrier is safe.
__device__ __forceinline__ void de_stamp(const DEArgs& g) {
  if (!g.done_ts) return;  // kernel argument: uniform
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t n = gridDim.x * gridDim.y;
    if (atomicAdd(g.tiles_done, 1u) + 1u == n) {
      const unsigned long long t = wall_clock64();
      __hip_atomic_store(g.done_ts, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Splat weights of every radius 0..R, computed once per workgroup with
// de_weight's arithmetic: entry (i + r) * (2r + 1) + (j + r) of radius r
// starts at de_wt_off(r) = sum over q < r of (2q + 1)^2.
__device__ __forceinline__ int de_wt_off(int r) { return (4 * r * r * r - r) / 3; }

__global__ __launch_bounds__(DE_TW * DE_TH) void k_de_tiled(DEArgs g) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int R = g.R;
  const int SW = DE_TW + 2 * R, SH = DE_TH + 2 * R;
  float4* s_col = reinterpret_cast<float4*>(smem);
  // s_m[d * SH + row]: bit c set when the source in tile column c of that row
  // has a radius of at least d (SW <= 64 columns for R <= 16)
  unsigned long long* s_m = reinterpret_cast<unsigned long long*>(smem + sizeof(float4) * SW * SH);
  int* s_r = reinterpret_cast<int*>(s_m + (R + 1) * SH);  // source radius, -1: empty
  float* s_wt = reinterpret_cast<float*>(s_r + SW * SH);
  __shared__ int s_rmax;
  const int tx0 = blockIdx.x * DE_TW, ty0 = blockIdx.y * DE_TH;
  const int tid = threadIdx.x;
  if (tid == 0) s_rmax = 0;
  for (int r = 0; r <= R; ++r) {
    const int S = 2 * r + 1;
    DESrc m;
    m.r = (float)r;
    m.e = rf_div_rn(1.0f, (float)S);
    m.nrm = (r > 0) ? rf_div_rn(0.63661977236f, (float)(r * r)) : 1.0f;
    for (int q = tid; q < S * S; q += DE_TW * DE_TH) s_wt[de_wt_off(r) + q] = de_weight(m, q / S - r, q % S - r);
  }
  for (int q = tid; q < (R + 1) * SH; q += DE_TW * DE_TH) s_m[q] = 0ull;
  __syncthreads();
  int rloc = 0;
  for (int k = tid; k < SW * SH; k += DE_TW * DE_TH) {
    const int ly = k / SW, lx = k % SW;
    float4 col; DESrc m;
    de_source(g, tx0 - R + lx, ty0 - R + ly, col, m);
    s_col[k] = col;
    s_r[k] = (int)m.r;
    rloc = max(rloc, (int)m.r);
    for (int d = 0; d <= (int)m.r; ++d) atomicOr(&s_m[d * SH + ly], 1ull << lx);
  }
  // A source contributes only within its own radius (de_weight is 0 beyond
  // it), so offsets beyond the largest radius staged for this tile add
  // nothing: the loops stop there, and per offset only sources whose radius
  // reaches it are read.  The sum is unchanged bit for bit (zero-weight terms
  // were skipped before too).
  rloc = max(rloc, __shfl_xor(rloc, 32));
  rloc = max(rloc, __shfl_xor(rloc, 16));
  rloc = max(rloc, __shfl_xor(rloc, 8));
  rloc = max(rloc, __shfl_xor(rloc, 4));
  rloc = max(rloc, __shfl_xor(rloc, 2));
  rloc = max(rloc, __shfl_xor(rloc, 1));
  if ((tid & 63) == 0) atomicMax(&s_rmax, rloc);
  __syncthreads();
  const int RM = min(R, s_rmax);
  const int lx = tid % DE_TW, ly = tid / DE_TW;
  const int tx = tx0 + lx, ty = ty0 + ly;
  if (tx < (int)g.W && ty < (int)g.H) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    // Only sources whose radius reaches the row offset are visited (the row's
    // mask for |i|), in the order of the full (i, j) sweep: j ascending is
    // tile column descending, so the highest set bit comes first.  The terms
    // left out are the zero ones the sweep skipped, so the sum is the same.
    const int c_lo = lx + R - RM;  // tile columns c_lo .. c_lo + 2 RM, j = lx + R - c
    const unsigned long long wmask = (2ull << (2 * RM)) - 1ull;
    for (int i = -RM; i <= RM; ++i) {
      // source row cy = ty - i -> LDS row ly + R - i
      const int rowi = ly + R - i;
      const int row = rowi * SW;
      const int ai = i < 0 ? -i : i;
      unsigned long long wm = (s_m[ai * SH + rowi] >> c_lo) & wmask;
      while (wm) {
        const int b = 63 - __clzll((long long)wm);
        wm &= ~(1ull << b);
        const int k = row + c_lo + b;
        const int j = RM - b;
        const int r = s_r[k];
        const int aj = j < 0 ? -j : j;
        if (r >= aj) {
          const float wgt = s_wt[de_wt_off(r) + (i + r) * (2 * r + 1) + (j + r)];
          if (wgt != 0.0f) {
            const float4 c = s_col[k];
            acc.x += c.x * wgt; acc.y += c.y * wgt; acc.z += c.z * wgt; acc.w += c.w * wgt;
          }
        }
      }
    }
    const size_t o = (size_t)ty * g.W + tx;
    if (g.de_out) g.de_out[o] = acc;
    if (g.img_out && g.do_tonemap) g.img_out[o] = tonemap_px(acc, g.gamma, g.brightness, g.vibrancy, g.k);
  }
  de_stamp(g);
}

// Radii beyond the LDS tile: the same gather straight from memory.
__global__ __launch_bounds__(256) void k_de_global(DEArgs g) {
  const int tx = blockIdx.x * 16 + (threadIdx.x % 16);
  const int ty = blockIdx.y * 16 + (threadIdx.x / 16);
  if (tx < (int)g.W && ty < (int)g.H) {
    const int R = g.R;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = -R; i <= R; ++i) {
      for (int j = -R; j <= R; ++j) {
        float4 c; DESrc m;
        de_source(g, tx - j, ty - i, c, m);
        const float wgt = de_weight(m, i, j);
        if (wgt != 0.0f) { acc.x += c.x * wgt; acc.y += c.y * wgt; acc.z += c.z * wgt; acc.w += c.w * wgt; }
      }
    }
    const size_t o = (size_t)ty * g.W + tx;
    if (g.de_out) g.de_out[o] = acc;
    if (g.img_out && g.do_tonemap) g.img_out[o] = tonemap_px(acc, g.gamma, g.brightness, g.vibrancy, g.k);
  }
  de_stamp(g);
}

__global__ void k_tonemap(const float4* __restrict__ in, float4* __restrict__ out, uint32_t n,
                          float gamma, float brightness, float vib, float k) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = tonemap_px(in[i], gamma, brightness, vib, k);
}

// RGBA float -> RGBA8: clamp to [0, 1], x255, round to nearest (NaN -> 0);
// one thread per pixel, 16 bytes in, 4 bytes out.
__global__ void k_rgba8(const float4* __restrict__ in, uchar4* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = in[i];
  const float c[4] = {v.x, v.y, v.z, v.w};
  uint8_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float t = c[k] > 0.0f ? (c[k] < 1.0f ? c[k] : 1.0f) : 0.0f;
    o[k] = (uint8_t)(t * 255.0f + 0.5f);
  }
  out[i] = make_uchar4(o[0], o[1], o[2], o[3]);
}

// ------------------------------------------------------------ diagnostics
__global__ void k_eval_xform(const Block* __restrict__ blk, const float* __restrict__ pts,
                             const float* __restrict__ rands, float* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TableRng rng{rands + (size_t)i * 16, 0};
  const XOut o = apply_xform(blk, 0, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], false, rng);
  out[4 * i + 0] = o.x;
  out[4 * i + 1] = o.y;
  out[4 * i + 2] = o.c;
  out[4 * i + 3] = o.w;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

template <int HM>
__global__ __launch_bounds__(256) void k_atomic_probe(float* bins, uint64_t nbins, uint32_t per_lane,
                                                      const float4* __restrict__ palette) {
  __shared__ float4 s_pal[256];
  s_pal[threadIdx.x] = palette[threadIdx.x];
  __syncthreads();
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  for (uint32_t k = 0; k < per_lane; ++k) {
    const uint32_t h = fmix32(t * 0x9E3779B9u ^ (k * 0x632BE5ABu));
    const uint32_t b = (uint32_t)(((uint64_t)h * nbins) >> 32);
    bin_sample<HM>(bins, b, true, h & 255u, 1.0f, s_pal);
  }
}

}  // namespace

__global__ void k_zero_u64(unsigned long long* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_animate(Block* blocks, uint32_t T, float tss_width, const float* rot, hipStream_t s) {
  hipLaunchKernelGGL(k_animate, dim3(T), dim3(256), 0, s, blocks, T, tss_width, rot);
  return hipGetLastError();
}

hipError_t launch_de_tonemap(const DEArgs& a, hipStream_t s) {
  const int SW = DE_TW + 2 * a.R, SH = DE_TH + 2 * a.R;
  const size_t n_wt = (size_t)(4 * (a.R + 1) * (a.R + 1) * (a.R + 1) - (a.R + 1)) / 3;  // de_wt_off(R + 1)
  const size_t sm = (size_t)SW * SH * (sizeof(float4) + sizeof(int)) + (size_t)(a.R + 1) * SH * 8 +
                    n_wt * sizeof(float);
  if (sm <= 65536) {  // within the default dynamic-LDS limit of one launch (R <= 14)
    dim3 grid((a.W + DE_TW - 1) / DE_TW, (a.H + DE_TH - 1) / DE_TH);
    hipLaunchKernelGGL(k_de_tiled, grid, dim3(DE_TW * DE_TH), sm, s, a);
  } else {
    dim3 grid((a.W + 15) / 16, (a.H + 15) / 16);
    hipLaunchKernelGGL(k_de_global, grid, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_tonemap(const float4* in, float4* out, uint32_t n, float gamma, float brightness,
                          float vibrancy, float k, hipStream_t s) {
  hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n, gamma, brightness,
                     vibrancy, k);
  return hipGetLastError();
}

hipError_t launch_rgba8(const float4* in, uchar4* out, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_rgba8, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_eval_xform(const Block* blk, const float* pts, const float* rands, float* out,
                             uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_eval_xform, dim3((n + 255) / 256), dim3(256), 0, s, blk, pts, rands, out, n);
  return hipGetLastError();
}

hipError_t launch_atomic_probe(float* bins, uint64_t nbins, uint64_t samples, int hm,
                               const float4* palette, hipStream_t s) {
  const uint32_t per_lane = 64;
  const uint64_t lanes = samples / per_lane;
  dim3 grid((uint32_t)(lanes / 256u)), block(256);
  switch (hm) {
    case 2: hipLaunchKernelGGL(k_atomic_probe<2>, grid, block, 0, s, bins, nbins, per_lane, palette); break;
    case 3: hipLaunchKernelGGL(k_atomic_probe<3>, grid, block, 0, s, bins, nbins, per_lane, palette); break;
    default: hipLaunchKernelGGL(k_atomic_probe<1>, grid, block, 0, s, bins, nbins, per_lane, palette); break;
  }
  return hipGetLastError();
}

}  // namespace rfk
