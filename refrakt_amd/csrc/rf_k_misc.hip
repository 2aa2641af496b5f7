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
// A splat reaches span = max(R, rmin) texels from its centre.

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

// Splat weights of every radius 0..R with de_weight's arithmetic: entry
// (i + r) * (2r + 1) + (r - j) of radius r (each row's columns in descending
// order, the order in which the gather walks a chord) starts at
// de_wt_off(r) = sum over q < r of (2q + 1)^2.
__device__ __forceinline__ int de_wt_off(int r) { return (4 * r * r * r - r) / 3; }

__global__ __launch_bounds__(256) void k_de_weights(float* __restrict__ wt, int R) {
  for (int r = 0; r <= R; ++r) {
    const int S = 2 * r + 1;
    DESrc m;
    m.r = (float)r;
    m.e = rf_div_rn(1.0f, (float)S);
    m.nrm = (r > 0) ? rf_div_rn(0.63661977236f, (float)(r * r)) : 1.0f;
    for (int q = threadIdx.x; q < S * S; q += blockDim.x) wt[de_wt_off(r) + q] = de_weight(m, q / S - r, r - q % S);
  }
}

// Every bin's splat radius (de_source's arithmetic), -1 for an empty bin.
__global__ __launch_bounds__(256) void k_de_radius(DEArgs g) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)g.W * g.H) return;
  const float a = g.hist[i].w;
  int r = -1;
  if (a != 0.0f) {
    const float q = rf_div_rn((float)g.R, powf(a, g.curve));
    int ri = (q >= (float)g.R) ? g.R : (int)q;
    if (!(q == q)) ri = g.R;
    r = ri > g.rmin ? ri : g.rmin;
    if (r < 0) r = 0;
  }
  g.rad[i] = (signed char)r;
}

// A TW x TH tile of texels per workgroup of 256 threads (TH / 8 texels per
// thread) from an LDS copy of the (TW + 2R) x (TH + 2R) sources around it.
// Per source row the tile keeps one 64-bit column mask per radius r (the
// sources whose radius is exactly r) and a mask of the radii present, so a
// texel visits, for row offset i, only the radii r >= |i| present in that
// row and, per radius, only the columns within r of its own: every source
// it reads contributes (up to the disc test inside the square).
constexpr int DE_TW = 32;
// Row strides of the per-radius tables, powers of two so that indexing by
// radius is a shift (a multiply by 2R + 1 or by the tile height is a
// quarter-rate integer instruction in the innermost (row, radius) loop):
// s_e rows of DE_ES >= TH + 2R entries, s_jx rows of DE_JS >= 2R + 1.
constexpr int DE_ES_SHIFT = 6, DE_JS_SHIFT = 5;
constexpr int DE_ES = 1 << DE_ES_SHIFT, DE_JS = 1 << DE_JS_SHIFT;

template <int TH, int NTH>
__global__ __launch_bounds__(NTH) void k_de_tiled(DEArgs g) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int R = max(g.R, g.rmin);  // span
  const int SW = DE_TW + 2 * R, SH = TH + 2 * R;
  float4* s_col = reinterpret_cast<float4*>(smem);
  unsigned long long* s_e = reinterpret_cast<unsigned long long*>(smem + sizeof(float4) * SW * SH);  // [r][row]
  uint32_t* s_rp = reinterpret_cast<uint32_t*>(s_e + (R + 1) * DE_ES);                                 // [row]
  float* s_wt = reinterpret_cast<float*>(s_rp + SH);
  // s_jx[r * DE_JS + i + R]: the columns j with a non-zero weight in row
  // offset i of radius r form the interval [lo, hi] (the disc's chord);
  // .x = (R - hi) | (index of w(i, hi) in s_wt) << 8 (the mask shift
  // relative to the texel's column, then the weights), .y = the chord's bit
  // mask (2 << (hi - lo)) - 1, 0 when the row has none.  One 8-byte entry,
  // read beside the source mask it applies to, so neither read waits on the
  // other; 8 bytes keep three workgroups per CU at R = 9.
  const size_t jx_off = ((size_t)(reinterpret_cast<unsigned char*>(s_wt + de_wt_off(R + 1)) - smem) + 7) & ~(size_t)7;
  uint2* s_jx = reinterpret_cast<uint2*>(smem + jx_off);
  __shared__ int s_rmax;
  const int tx0 = blockIdx.x * DE_TW, ty0 = blockIdx.y * TH;
  const int tid = threadIdx.x;
  if (tid == 0) s_rmax = 0;
  for (int q = tid; q < de_wt_off(R + 1); q += NTH) s_wt[q] = g.wt[q];
  for (int q = tid; q < (R + 1) * DE_ES; q += NTH) s_e[q] = 0ull;
  for (int q = tid; q < SH; q += NTH) s_rp[q] = 0u;
  for (int q = tid; q < (R + 1) * DE_JS; q += NTH) {
    const int r = q >> DE_JS_SHIFT, i = (q & (DE_JS - 1)) - R;
    uint2 v = make_uint2(0u, 0u);
    if (i >= -r && i <= r) {
      const float* w = g.wt + de_wt_off(r) + (i + r) * (2 * r + 1);
      int lo = r + 1, hi = -r - 1;
      for (int j = -r; j <= r; ++j)
        if (w[r - j] != 0.0f) { lo = min(lo, j); hi = max(hi, j); }
      // a chord is at most 2r + 1 <= 31 columns wide: its mask fits 32 bits
      if (lo <= hi)
        v = make_uint2((uint32_t)(R - hi) | ((uint32_t)(de_wt_off(r) + (i + r) * (2 * r + 1) + r - hi) << 8),
                       (2u << (hi - lo)) - 1u);
    }
    s_jx[q] = v;
  }
  __syncthreads();
  int rloc = 0;
  for (int k = tid; k < SW * SH; k += NTH) {
    const int ly = k / SW, lx = k % SW;
    const int cx = tx0 - R + lx, cy = ty0 - R + ly;
    const int px = cx + 1, py = (int)g.H - 1 - cy;
    if (cx < 0 || px >= (int)g.W || cy < 0 || cy >= (int)g.H) continue;
    const size_t b = (size_t)py * g.W + px;
    const int r = g.rad[b];
    if (r < 0) continue;
    s_col[k] = g.hist[b];
    atomicOr(&s_e[(r << DE_ES_SHIFT) + ly], 1ull << lx);
    atomicOr(&s_rp[ly], 1u << r);
    rloc = max(rloc, r);
  }
  rloc = max(rloc, __shfl_xor(rloc, 32));
  rloc = max(rloc, __shfl_xor(rloc, 16));
  rloc = max(rloc, __shfl_xor(rloc, 8));
  rloc = max(rloc, __shfl_xor(rloc, 4));
  rloc = max(rloc, __shfl_xor(rloc, 2));
  rloc = max(rloc, __shfl_xor(rloc, 1));
  if ((tid & 63) == 0) atomicMax(&s_rmax, rloc);
  __syncthreads();
  const int RM = min(R, s_rmax);
  const int lx = tid % DE_TW;
#pragma unroll 1
  for (int ly = tid / DE_TW; ly < TH; ly += NTH / DE_TW) {
    const int tx = tx0 + lx, ty = ty0 + ly;
    if (tx >= (int)g.W || ty >= (int)g.H) continue;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    // source (row ly + R - i, column lx + R - j) adds weight w_r(i, j)
    for (int i = -RM; i <= RM; ++i) {
      const int rowi = ly + R - i;
      const int ai = i < 0 ? -i : i;
      uint32_t radii = s_rp[rowi] >> ai;
      while (radii) {
        const int r = ai + __builtin_ctz(radii);
        radii &= radii - 1u;
        const uint2 jx = s_jx[(r << DE_JS_SHIFT) + i + R];
        const unsigned long long e = s_e[(r << DE_ES_SHIFT) + rowi];
        const int c0 = lx + (int)(jx.x & 0xffu);  // columns c0 .. c0 + hi - lo: j = hi .. lo
        uint32_t m = (uint32_t)(e >> c0) & jx.y;
        const float* w = s_wt + (jx.x >> 8);  // w[b]: j = hi - b
        const int base = rowi * SW + c0;
        while (m) {
          const int b = __builtin_ctz(m);
          m &= m - 1u;
          const float wgt = w[b];
          const float4 c = s_col[base + b];
          acc.x = fmaf(c.x, wgt, acc.x);
          acc.y = fmaf(c.y, wgt, acc.y);
          acc.z = fmaf(c.z, wgt, acc.z);
          acc.w = fmaf(c.w, wgt, acc.w);
        }
      }
    }
    const size_t o = (size_t)ty * g.W + tx;
    if (g.de_out) g.de_out[o] = acc;
    if (g.img_out && g.do_tonemap) g.img_out[o] = tonemap_px(acc, g.gamma, g.brightness, g.vibrancy, g.k);
  }
  de_stamp(g);
}

template <int TH>
size_t de_tiled_lds(int R) {
  const size_t SW = DE_TW + 2 * R, SH = TH + 2 * R;
  return SW * SH * sizeof(float4) + (size_t)(R + 1) * DE_ES * 8 + SH * 4 + de_weight_bytes(R) + 8 +
         (size_t)(R + 1) * DE_JS * 8;
}

// Radii beyond the LDS tile: the same gather straight from memory.
__global__ __launch_bounds__(256) void k_de_global(DEArgs g) {
  const int tx = blockIdx.x * 16 + (threadIdx.x % 16);
  const int ty = blockIdx.y * 16 + (threadIdx.x / 16);
  if (tx < (int)g.W && ty < (int)g.H) {
    const int R = max(g.R, g.rmin);
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
  if (!a.rad || !a.wt) return hipErrorInvalidValue;
  const uint64_t n = (uint64_t)a.W * a.H;
  // Tiles of 32 x 32 texels while they fit the 64 KiB of LDS one launch gets
  // by default (R <= 11), 32 x 8 up to R = 14 (64-column masks need
  // 32 + 2R <= 64 as well), the global-memory gather beyond.
  const int span = a.R > a.rmin ? a.R : a.rmin;
  // (chords of at most 31 columns keep the tap masks in 32 bits)
  // span <= 15 keeps 2 span + 1 <= DE_JS and TH + 2 span <= DE_ES
  const bool tall = de_tiled_lds<32>(span) <= 65536 && span <= 15;
  const bool flat = de_tiled_lds<8>(span) <= 65536 && 32 + 2 * span <= 64 && span <= 15;
  if (tall || flat) {
    hipLaunchKernelGGL(k_de_weights, dim3(1), dim3(256), 0, s, a.wt, span);
    if (n) hipLaunchKernelGGL(k_de_radius, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  }
  if (tall) {
    dim3 grid((a.W + DE_TW - 1) / DE_TW, (a.H + 31) / 32);
    // eight waves per 32 x 32 tile: the LDS allows three tiles per CU, so
    // 24 waves per CU hide the gather's LDS latency (12 with four waves)
    hipLaunchKernelGGL((k_de_tiled<32, 512>), grid, dim3(512), de_tiled_lds<32>(span), s, a);
  } else if (flat) {
    dim3 grid((a.W + DE_TW - 1) / DE_TW, (a.H + 7) / 8);
    hipLaunchKernelGGL((k_de_tiled<8, 256>), grid, dim3(256), de_tiled_lds<8>(span), s, a);
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
