// Denoiser network kernels (gfx950): bf16 MFMA convolutions with fp32
// accumulation, weight gradients, pooling, loss and Adam.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rf_nn.hpp"

namespace rfn {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even; NaN stays NaN.
__device__ __forceinline__ bf16 f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)0x7fc0u;
  return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// ------------------------------------------------------------ convolution
// Implicit GEMM: M = pixels, N = output channels, K = 9 * Cin.  A workgroup
// of 4 waves covers 64 consecutive pixels of one image row (16 per wave) and
// 16*NT output channels.  Per (r, s, 16-channel step) each wave issues NT
// v_mfma_f32_16x16x16_bf16 with one A fragment (16 pixels x 16 channels) and
// NT B fragments (16 channels x 16 outputs).  Operand layout of that
// instruction (wave64): lane l holds A[l % 16][4 (l / 16) + e] and
// B[4 (l / 16) + e][l % 16] for e = 0..3, and D[4 (l / 16) + e][l % 16].
// NHWC activations and [r][s][Cout][Cin] weights make both fragments one
// 8-byte load per lane.
template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(View in, const bf16* __restrict__ w, const float* __restrict__ bias,
                                                 View out, int H, int W, int Cin, int Cout, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.y / H, y = blockIdx.y % H;
  const int x0 = blockIdx.x * 64 + wave * 16;
  const int co0 = blockIdx.z * 16 * NT;
  const int i = lane & 15, kq = lane >> 4;
  float4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
  const int xa = x0 + i;
  for (int r = 0; r < 3; ++r) {
    const int yy = y + r - 1;
    const bool rok = yy >= 0 && yy < H;
    for (int s = 0; s < 3; ++s) {
      const int xx = xa + s - 1;
      const bool ok = rok && xx >= 0 && xx < W;
      const bf16* ap = in.p + ((size_t)(n * H + (rok ? yy : 0)) * W + (ok ? xx : 0)) * in.cs + in.co + 4 * kq;
      const bf16* wp = w + ((size_t)(r * 3 + s) * Cout + co0 + i) * Cin + 4 * kq;
      for (int c0 = 0; c0 < Cin; c0 += 16) {
        short4_t a = short4_t{0, 0, 0, 0};
        if (ok) a = *reinterpret_cast<const short4_t*>(ap + c0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const short4_t b = *reinterpret_cast<const short4_t*>(wp + (size_t)t * 16 * Cin + c0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = co0 + 16 * t + i;
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int px = x0 + 4 * kq + e;
      if (px < W) {
        float v = acc[t][e] + bv;
        if (relu) v = fmaxf(v, 0.0f);
        out.p[((size_t)(n * H + y) * W + px) * out.cs + out.co + co] = f2bf(v);
      }
    }
  }
}

__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

// LDS-tiled implicit GEMM.  A workgroup of 4 waves covers 4 RW image rows
// x CT_W = 64 pixels (RW rows per wave, four 16-pixel A fragments per row) and
// 16*NT output channels.  Per KC-channel step of Cin (KC = 32 with gfx950's
// v_mfma_f32_16x16x32_bf16 for the wide layers, else 16 with the 16x16x16 form)
// it stages the (4 RW + 2) x (CT_W + 2) input patch (zero halo) and the
// step's 3x3 x 16*NT x KC weights in LDS, KC + 8 channels per pixel / weight
// row (the pad spreads 16 consecutive rows over distinct banks and keeps
// 16-byte stores aligned); every A fragment then feeds NT MFMAs and every B
// fragment four per output row it reaches.  With RW = 2 a wave's four input
// rows serve six (row, tap row) pairs, so each A fragment read from LDS feeds
// up to 2 NT MFMAs: the narrow layers, whose NT is small, read a third fewer
// A fragments and half the B fragments per output.  The next step's operands
// are loaded into registers while the MFMAs of the current one run.  Needs 16-byte aligned channel runs (Cin,
// channel strides and offsets multiples of 8).
// Operand layouts (wave64, i = lane % 16, kq = lane / 16): 16x16x16 lane
// holds A[i][4 kq + e], B[4 kq + e][i] (e < 4); 16x16x32 lane holds
// A[i][8 kq + e], B[8 kq + e][i] (e < 8); both give D[4 kq + e][i] (e < 4).
constexpr int CT_W = 64;
// staged input pixels of a tile of 4 RW rows
__host__ __device__ constexpr int ct_in_pix(int RW) { return (4 * RW + 2) * (CT_W + 2); }

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <int NT, int KC, int RW>
struct ConvStage {
  static constexpr int PS = KC + 8;             // LDS row stride (bf16)
  static constexpr int PQ = KC / 8;             // 16-byte pieces per pixel / weight row
  static constexpr int IN_Q = ct_in_pix(RW) * PQ;
  static constexpr int W_Q = 9 * 16 * NT * PQ;
  static constexpr int IN_PER = (IN_Q + 255) / 256;
  static constexpr int W_PER = (W_Q + 255) / 256;
  uint4 vin[IN_PER];
  uint4 vw[W_PER];
};

// Channels [0, Cup) come from the half-resolution `up` (nearest-neighbour
// upsampling read in place), the rest from `in`; a KC step never straddles
// the two (Cup is a multiple of 32).
template <int NT, int KC, int RW>
__device__ __forceinline__ void conv_load(ConvStage<NT, KC, RW>& st, const View& in, const View& up, int Cup,
                                          const bf16* __restrict__ w, int n, int y0, int x0, int co0, int c0, int H,
                                          int W, int Cin, int Cout) {
  using S = ConvStage<NT, KC, RW>;
  const int tid = threadIdx.x;
  const bool from_up = c0 < Cup;  // workgroup-uniform
#pragma unroll
  for (int k = 0; k < S::IN_PER; ++k) {
    const int q = tid + 256 * k;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (q < S::IN_Q) {
      const int piece = q % S::PQ, pix = q / S::PQ;
      const int yy = y0 + pix / (CT_W + 2) - 1, gx = x0 + pix % (CT_W + 2) - 1;
      if (yy >= 0 && yy < H && gx >= 0 && gx < W) {
        const bf16* src = from_up ? up.p + ((size_t)(n * (H >> 1) + (yy >> 1)) * (W >> 1) + (gx >> 1)) * up.cs +
                                        up.co + c0 + 8 * piece
                                  : in.p + ((size_t)(n * H + yy) * W + gx) * in.cs + in.co + (c0 - Cup) + 8 * piece;
        v = *reinterpret_cast<const uint4*>(src);
      }
    }
    st.vin[k] = v;
  }
#pragma unroll
  for (int k = 0; k < S::W_PER; ++k) {
    const int q = tid + 256 * k;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (q < S::W_Q) {
      const int piece = q % S::PQ, row = q / S::PQ;  // row = tap * 16 * NT + output channel
      const int tap = row / (16 * NT), co = row % (16 * NT);
      v = *reinterpret_cast<const uint4*>(w + ((size_t)tap * Cout + co0 + co) * Cin + c0 + 8 * piece);
    }
    st.vw[k] = v;
  }
}

// One lane's operand fragment: 4 (16x16x16) or 8 (16x16x32) bf16 values.
template <int KC>
struct ConvFrag;
template <>
struct ConvFrag<16> {
  typedef short4_t T;
  static __device__ __forceinline__ float4_t mfma(T a, T b, float4_t acc) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
};
template <>
struct ConvFrag<32> {
  typedef bf16x8_t T;
  static __device__ __forceinline__ float4_t mfma(T a, T b, float4_t acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
};

// Output of one tile: bias, ReLU and bf16 in LDS (tile row RW * wave + rr is
// image row y0 + RW * wave + rr), then 16-byte stores per pixel, the fused
// 2x2 averages or the fused fp32 residual.  The output tile reuses the
// staging buffer: the caller's next step waits on a barrier before staging.
template <int NT, int RW>
__device__ __forceinline__ void conv_store_lds(const float4_t (&acc)[RW][4][NT], bf16* s_buf,
                                               const float* __restrict__ bias, const View& out, const ConvFuse& fz,
                                               int relu, int n, int y0, int x0, int co0, int H, int W) {
  constexpr int OC = 16 * NT, OS = OC + 8, TR = 4 * RW;
  constexpr int PQ = OC / 8;  // 16-byte pieces per pixel
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, kq = lane >> 4;
  __syncthreads();  // the last step's fragment reads are done: the staging buffer is free
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    bf16* const s_o = s_buf + (RW * wave + rr) * CT_W * OS;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const float bv = bias ? bias[co0 + 16 * t + i] : 0.0f;
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float v = acc[rr][f][t][e] + bv;
          if (relu) v = fmaxf(v, 0.0f);
          s_o[(16 * f + 4 * kq + e) * OS + 16 * t + i] = f2bf(v);
        }
    }
  }
  __syncthreads();
  if (fz.y) {  // workgroup-uniform: the residual output replaces `out`
    const int px = x0 + lane;
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int y = y0 + RW * wave + rr;
      if (y < H && px < W) {
        const size_t p = (size_t)(n * H + y) * W + px;
        const uint2 xv = *reinterpret_cast<const uint2*>(fz.res.p + p * fz.res.cs + fz.res.co);
        const uint2 ov = *reinterpret_cast<const uint2*>(s_buf + ((RW * wave + rr) * CT_W + lane) * OS);
        *reinterpret_cast<float4*>(fz.y + p * 4) =
            make_float4(bf_lo(xv.x) + bf_lo(ov.x), bf_hi(xv.x) + bf_hi(ov.x), bf_lo(xv.y) + bf_lo(ov.y),
                        bf_hi(xv.y) + bf_hi(ov.y));
      }
    }
  } else {
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int y = y0 + RW * wave + rr;
      const bf16* const s_o = s_buf + (RW * wave + rr) * CT_W * OS;
      if (y < H) {
#pragma unroll
        for (int q = lane; q < CT_W * PQ; q += 64) {
          const int px = q / PQ, piece = q % PQ;
          if (x0 + px < W)
            *reinterpret_cast<uint4*>(out.p + ((size_t)(n * H + y) * W + x0 + px) * out.cs + out.co + co0 +
                                      8 * piece) = *reinterpret_cast<const uint4*>(s_o + px * OS + 8 * piece);
        }
      }
    }
    if (fz.pool.p) {  // workgroup-uniform: the tile's 2x2 averages, summed in k_avgpool2_v8's order
      const View& pool = fz.pool;
      constexpr int PW = CT_W / 2, NPQ = (TR / 2) * PW * PQ;
      for (int q = tid; q < NPQ; q += 256) {
        const int piece = q % PQ, pp = (q / PQ) % PW, pr = q / (PQ * PW);
        const int yo = (y0 >> 1) + pr, xo = (x0 >> 1) + pp;
        if (yo < (H >> 1) && xo < (W >> 1)) {
          float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              const uint4 v =
                  *reinterpret_cast<const uint4*>(s_buf + (2 * pr + a) * CT_W * OS + (2 * pp + b) * OS + 8 * piece);
              const uint32_t qv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
              for (int k = 0; k < 4; ++k) { sum[2 * k] += bf_lo(qv[k]); sum[2 * k + 1] += bf_hi(qv[k]); }
            }
          uint32_t r[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            r[k] = (uint32_t)f2bf(0.25f * sum[2 * k]) | ((uint32_t)f2bf(0.25f * sum[2 * k + 1]) << 16);
          *reinterpret_cast<uint4*>(pool.p + ((size_t)(n * (H >> 1) + yo) * (W >> 1) + xo) * pool.cs + pool.co +
                                    co0 + 8 * piece) = make_uint4(r[0], r[1], r[2], r[3]);
        }
      }
    }
  }
}

// One workgroup per tile (x block, row block of 4 RW rows of image n, 16 NT
// output channel block): workgroup t has channel block t % nbz, x block
// (t / nbz) % nbx and row block t / (nbz nbx), so the workgroups that read
// the same input patch for different output channels run next to each other.
// (A persistent variant, each workgroup walking tiles b, b + G, ... with the
// next tile's first step prefetched during the last step of the current one,
// measured slower: 16.8 ms against 14.0 ms per 7680x4320 inference, its
// registers no longer fitting three waves per SIMD without spills.)
template <int NT, int KC, int RW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RW == 2 ? (KC == 16 ? 3 : 2) : 1))) void k_conv3x3_lds(
    View in, const bf16* __restrict__ w, const float* __restrict__ bias, View out, int H, int W, int Cin, int Cout,
    int relu, int vec_out, ConvFuse fz, int nbx, int nbz, int n_tiles) {
  const View& up = fz.up;
  const int Cup = fz.Cup;
  using S = ConvStage<NT, KC, RW>;
  using F = ConvFrag<KC>;
  constexpr int PS = S::PS, KE = KC / 4;  // fragment elements per lane
  constexpr int OC = 16 * NT, OS = OC + 8;  // output tile: channels per pixel, LDS row stride
  constexpr int TR = 4 * RW;                // tile rows
  constexpr int IN_PIX = ct_in_pix(RW);
  static_assert(TR * CT_W * OS <= IN_PIX * PS + 9 * 16 * NT * PS, "output tile fits the staging buffer");
  __shared__ __attribute__((aligned(16))) bf16 s_buf[IN_PIX * PS + 9 * 16 * NT * PS];
  bf16* const s_in = s_buf;
  bf16* const s_w = s_buf + IN_PIX * PS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nrb = (H + TR - 1) / TR;
  const int i = lane & 15, kq = lane >> 4;
  const int tile = blockIdx.x;
  if (tile >= n_tiles) return;  // workgroup-uniform
  const int bz = tile % nbz, rest = tile / nbz;
  const int bx = rest % nbx, by = rest / nbx;
  const int n = by / nrb, y0 = (by % nrb) * TR, x0 = bx * CT_W, co0 = bz * 16 * NT;
  float4_t acc[RW][4][NT];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr)
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[rr][f][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  S st;
  conv_load<NT, KC, RW>(st, in, up, Cup, w, n, y0, x0, co0, 0, H, W, Cin, Cout);
  for (int c0 = 0; c0 < Cin; c0 += KC) {
    __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
    for (int k = 0; k < S::IN_PER; ++k) {
      const int q = tid + 256 * k;
      if (q < S::IN_Q) *reinterpret_cast<uint4*>(s_in + (q / S::PQ) * PS + 8 * (q % S::PQ)) = st.vin[k];
    }
#pragma unroll
    for (int k = 0; k < S::W_PER; ++k) {
      const int q = tid + 256 * k;
      if (q < S::W_Q) *reinterpret_cast<uint4*>(s_w + (q / S::PQ) * PS + 8 * (q % S::PQ)) = st.vw[k];
    }
    __syncthreads();
    if (c0 + KC < Cin) conv_load<NT, KC, RW>(st, in, up, Cup, w, n, y0, x0, co0, c0 + KC, H, W, Cin, Cout);
    if constexpr (RW == 1) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const bf16* bp = s_w + ((r * 3 + s) * 16 * NT + i) * PS + KE * kq;
          const bf16* ap = s_in + ((wave + r) * (CT_W + 2) + i + s) * PS + KE * kq;
#pragma unroll
          for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int t = 0; t < NT; ++t)
              acc[0][f][t] = F::mfma(*reinterpret_cast<const typename F::T*>(ap + 16 * f * PS),
                                     *reinterpret_cast<const typename F::T*>(bp + 16 * t * PS), acc[0][f][t]);
        }
    } else {
      // Staged input row j of the wave's RW + 2 rows is tap row r = j - rr
      // of its output row rr: per tap column s, the 3 NT B fragments are
      // read once and every A fragment once, for all the rows it reaches.
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        typename F::T b[3][NT];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            b[r][t] = *reinterpret_cast<const typename F::T*>(s_w + ((r * 3 + s) * 16 * NT + 16 * t + i) * PS +
                                                              KE * kq);
#pragma unroll
        for (int j = 0; j < RW + 2; ++j) {
          const bf16* ap = s_in + ((RW * wave + j) * (CT_W + 2) + i + s) * PS + KE * kq;
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const typename F::T a = *reinterpret_cast<const typename F::T*>(ap + 16 * f * PS);
#pragma unroll
            for (int rr = 0; rr < RW; ++rr) {
              const int r = j - rr;
              if (r >= 0 && r < 3) {
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[rr][f][t] = F::mfma(a, b[r][t], acc[rr][f][t]);
              }
            }
          }
        }
      }
    }
  }
  if (vec_out) {  // workgroup-uniform
    conv_store_lds<NT, RW>(acc, s_buf, bias, out, fz, relu, n, y0, x0, co0, H, W);
  } else {
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int y = y0 + RW * wave + rr;
      if (y < H) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int co = co0 + 16 * t + i;
          const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
          for (int f = 0; f < 4; ++f) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int px = x0 + 16 * f + 4 * kq + e;
              if (px < W) {
                float v = acc[rr][f][t][e] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                out.p[((size_t)(n * H + y) * W + px) * out.cs + out.co + co] = f2bf(v);
              }
            }
          }
        }
      }
    }
  }
}


// Weight gradient on MFMA: an implicit GEMM per tap (r, s) with M = output
// channels, N = input channels, K = pixels:
//   dw[r][s][k][c] = sum_p dy[p][k] x[p + (r - 1, s - 1)][c].
// A workgroup owns a TCO x TCI channel tile, all nine taps, and a run of
// "units" (one 64-pixel strip of one dy row; units are strip-major, so a
// run walks down the rows of a strip).  The three input rows around the
// current dy row live in a ring of LDS slots (image row yy in slot yy % 3):
// moving one row down loads one new input row, a new image or strip two.
// Staging is transposed to channel-major ([channel][pixel]) so a lane's eight
// consecutive pixels are one 16-byte fragment read, and each input row is
// stored at the three column shifts s = 0, 1, 2 so every tap reads aligned
// fragments.  The next unit's rows are loaded into registers while the
// current unit's MFMAs run (v_mfma_f32_16x16x32_bf16: lane l holds
// A[l % 16][8 (l / 16) + e], B[8 (l / 16) + e][l % 16], D[4 (l / 16) + e][l % 16]).
// Waves own SPW 16-channel output blocks x one 16-channel input block; when
// the tile has fewer such groups than waves, WK = 2 waves split each strip's
// two 32-pixel K steps and add in LDS at the end.  The bias gradient is the
// product with an all-ones B.  Each run's sums go to its own slice of `part`
// (no atomics); k_wgrad_reduce adds the slices to dw and db in a fixed order.
constexpr int WG_SEG = 64, WG_XS = WG_SEG + 8;  // strip width, LDS row stride (bf16)

template <int TCO, int TCI, int SPW>
struct WgradCfg {
  static constexpr int GROUPS = (TCO / (16 * SPW)) * (TCI / 16);  // wave groups of the tile
  static constexpr int WK = GROUPS >= 4 ? 1 : 2;                  // waves per group (K split)
  static_assert(GROUPS * WK <= 4 && (GROUPS == 4 || WK == 2), "four waves cover the tile");
  static constexpr int PQO = TCO / 8, PQI = TCI / 8;               // 16-byte pieces per pixel
  static constexpr int DY_Q = WG_SEG * PQO, XROW_Q = (WG_SEG + 2) * PQI;
  static constexpr int NPER = (DY_Q + 3 * XROW_Q + 255) / 256;    // pieces per thread, at most 3 rows
  static constexpr int LDS_X = 3 * 3 * TCI * WG_XS, LDS_DY = TCO * WG_XS;  // bf16 elements
  static constexpr int RED = 10 * SPW * 4 * 64;                   // floats of the K-split reduction
  static_assert(WK == 1 || GROUPS * RED * 4 <= LDS_X * 2, "the reduction fits the input ring");
};

// What one stage loads: dy row y of image n in strip xs and the nr (at most
// three) consecutive input rows y0 .. y0 + nr - 1 of the same image and strip.
struct WgradStage {
  int n, y, xs, y0, nr;
};

__device__ __forceinline__ WgradStage wgrad_stage(int u, int H, int rows, bool fresh) {
  WgradStage st;
  const int strip = u / rows, row = u % rows;
  st.n = row / H;
  st.y = row % H;
  st.xs = strip * WG_SEG;
  // a run's first unit, a new image or a new strip: rows y - 1 .. y + 1;
  // otherwise only y + 1 is new
  const int lo = (fresh || st.y == 0) ? st.y - 1 : st.y + 1;
  st.y0 = max(lo, 0);
  st.nr = max(0, min(st.y + 1, H - 1) - st.y0 + 1);
  return st;
}

template <int TCO, int TCI, int SPW>
__device__ __forceinline__ void wgrad_load(uint4* v, const WgradStage& st,
                                           const View& x, const View& dy, int H, int W, int ci0, int co0) {
  using C = WgradCfg<TCO, TCI, SPW>;
  const int tid = threadIdx.x;
  const int total = C::DY_Q + st.nr * C::XROW_Q;
#pragma unroll
  for (int k = 0; k < C::NPER; ++k) {
    const int q = tid + 256 * k;
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    if (q < C::DY_Q) {
      const int p = q / C::PQO, pc = q % C::PQO;
      if (st.xs + p < W)
        r = *reinterpret_cast<const uint4*>(dy.p + ((size_t)(st.n * H + st.y) * W + st.xs + p) * dy.cs + dy.co +
                                            co0 + 8 * pc);
    } else if (q < total) {
      const int qq = q - C::DY_Q;
      const int ri = qq / C::XROW_Q, rem = qq % C::XROW_Q;
      const int p = rem / C::PQI, pc = rem % C::PQI;
      const int gx = st.xs + p - 1;
      if (gx >= 0 && gx < W)
        r = *reinterpret_cast<const uint4*>(x.p + ((size_t)(st.n * H + st.y0 + ri) * W + gx) * x.cs + x.co + ci0 +
                                            8 * pc);
    }
    v[k] = r;
  }
}

template <int TCO, int TCI, int SPW>
__device__ __forceinline__ void wgrad_store(const uint4* v, const WgradStage& st,
                                            bf16* s_x, bf16* s_dy) {
  using C = WgradCfg<TCO, TCI, SPW>;
  const int tid = threadIdx.x;
  const int total = C::DY_Q + st.nr * C::XROW_Q;
#pragma unroll
  for (int k = 0; k < C::NPER; ++k) {
    const int q = tid + 256 * k;
    const uint32_t w4[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
    if (q < C::DY_Q) {
      const int p = q / C::PQO, pc = q % C::PQO;
      bf16* d = s_dy + (8 * pc) * WG_XS + p;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        d[(2 * e) * WG_XS] = (bf16)(w4[e] & 0xffffu);
        d[(2 * e + 1) * WG_XS] = (bf16)(w4[e] >> 16);
      }
    } else if (q < total) {
      const int qq = q - C::DY_Q;
      const int ri = qq / C::XROW_Q, rem = qq % C::XROW_Q;
      const int p = rem / C::PQI, pc = rem % C::PQI;
      const int slot = (st.y0 + ri) % 3;
#pragma unroll
      for (int sh = 0; sh < 3; ++sh) {
        // copy sh holds x[xs + j + sh - 1] at column j; staged pixel p is x[xs + p - 1]
        const int j = p - sh;
        if (j < 0 || j >= WG_SEG) continue;
        bf16* d = s_x + ((slot * 3 + sh) * TCI + 8 * pc) * WG_XS + j;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          d[(2 * e) * WG_XS] = (bf16)(w4[e] & 0xffffu);
          d[(2 * e + 1) * WG_XS] = (bf16)(w4[e] >> 16);
        }
      }
    }
  }
}

template <int TCO, int TCI, int SPW>
__global__ __launch_bounds__(256) void k_wgrad_mfma(View x, View dy, float* __restrict__ part, int N, int H, int W,
                                                    int Cin, int Cout, int runs, int with_bias) {
  using C = WgradCfg<TCO, TCI, SPW>;
  __shared__ __attribute__((aligned(16))) bf16 s_x[C::LDS_X];   // [slot][shift][ci][px]
  __shared__ __attribute__((aligned(16))) bf16 s_dy[C::LDS_DY]; // [co][px]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, kq = lane >> 4;
  const int run = blockIdx.x, cb = blockIdx.y, kb = blockIdx.z;
  const int ci0 = cb * TCI, co0 = kb * TCO;
  const int rows = N * H, strips = (W + WG_SEG - 1) / WG_SEG, units = rows * strips;
  const int u0 = (int)((long long)units * run / runs), u1 = (int)((long long)units * (run + 1) / runs);
  // this wave's group: output blocks sco0 .. sco0 + SPW - 1, input block sci, K steps kw, kw + WK, ...
  const int kw = wave % C::WK, grp = wave / C::WK;
  const bool active = grp < C::GROUPS;
  const int sci = grp % (TCI / 16), sco0 = (grp / (TCI / 16)) * SPW;
  const bool bias = with_bias && cb == 0 && sci == 0 && active;
  float4_t acc[9][SPW], accb[SPW];
#pragma unroll
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int t = 0; t < SPW; ++t) acc[q][t] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < SPW; ++t) accb[t] = float4_t{0.f, 0.f, 0.f, 0.f};
  const uint4 one4 = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);
  const bf16x8_t ones = __builtin_bit_cast(bf16x8_t, one4);
  uint4 v[C::NPER];
  if (u0 < u1) {
    WgradStage st = wgrad_stage(u0, H, rows, true);
    wgrad_load<TCO, TCI, SPW>(v, st, x, dy, H, W, ci0, co0);
    wgrad_store<TCO, TCI, SPW>(v, st, s_x, s_dy);
    __syncthreads();
    for (int u = u0; u < u1; ++u) {
      const WgradStage cur = st;
      const bool more = u + 1 < u1;
      if (more) {
        st = wgrad_stage(u + 1, H, rows, false);
        wgrad_load<TCO, TCI, SPW>(v, st, x, dy, H, W, ci0, co0);
      }
      if (active) {
#pragma unroll
        for (int ks = kw; ks < WG_SEG / 32; ks += C::WK) {
          const int px = 32 * ks + 8 * kq;
          bf16x8_t a[SPW];
#pragma unroll
          for (int t = 0; t < SPW; ++t)
            a[t] = *reinterpret_cast<const bf16x8_t*>(s_dy + (16 * (sco0 + t) + i) * WG_XS + px);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const int yy = cur.y + r - 1;
            if (yy < 0 || yy >= H) continue;  // zero padding: no products
            const int slot = yy % 3;
#pragma unroll
            for (int sh = 0; sh < 3; ++sh) {
              const bf16x8_t b =
                  *reinterpret_cast<const bf16x8_t*>(s_x + ((slot * 3 + sh) * TCI + 16 * sci + i) * WG_XS + px);
#pragma unroll
              for (int t = 0; t < SPW; ++t)
                acc[r * 3 + sh][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b, acc[r * 3 + sh][t], 0, 0, 0);
            }
          }
          if (bias) {
#pragma unroll
            for (int t = 0; t < SPW; ++t) accb[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], ones, accb[t], 0, 0, 0);
          }
        }
      }
      __syncthreads();  // the ring slot and dy buffer the next stage overwrites are read
      if (more) {
        wgrad_store<TCO, TCI, SPW>(v, st, s_x, s_dy);
        __syncthreads();
      }
    }
  }
  // K-split groups: the second wave of each group adds its sums to the first's
  if (C::WK == 2) {
    float* s_red = reinterpret_cast<float*>(s_x);  // free: the last unit's reads ended at a barrier
    if (active && kw == 1) {
#pragma unroll
      for (int q = 0; q < 10; ++q)
#pragma unroll
        for (int t = 0; t < SPW; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            s_red[((grp * 10 + q) * SPW + t) * 4 * 64 + e * 64 + lane] = q < 9 ? acc[q][t][e] : accb[t][e];
    }
    __syncthreads();
    if (active && kw == 0) {
#pragma unroll
      for (int q = 0; q < 10; ++q)
#pragma unroll
        for (int t = 0; t < SPW; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float o = s_red[((grp * 10 + q) * SPW + t) * 4 * 64 + e * 64 + lane];
            if (q < 9) acc[q][t][e] += o;
            else accb[t][e] += o;
          }
    }
  }
  if (!active || kw != 0) return;
  // part[run][q][co][ci]; the bias sums of every run follow all runs' weights
  const size_t per = (size_t)9 * Cout * Cin;
  float* const pr = part + (size_t)run * per;
#pragma unroll
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int t = 0; t < SPW; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int co = co0 + 16 * (sco0 + t) + 4 * kq + e, ci = ci0 + 16 * sci + i;
        pr[((size_t)q * Cout + co) * Cin + ci] = acc[q][t][e];
      }
  if (bias && i == 0) {
    // every column of the all-ones product holds the row sums
#pragma unroll
    for (int t = 0; t < SPW; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        part[(size_t)runs * per + (size_t)run * Cout + co0 + 16 * (sco0 + t) + 4 * kq + e] = accb[t][e];
  }
}

// dw[i] += sum over runs of part (bias entries after the weights): 64
// outputs per workgroup, 16 lanes per output summing every 16th run, the 16
// sums added in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_wgrad_reduce(const float* __restrict__ part, float* __restrict__ dw,
                                                       float* __restrict__ db, int runs, int n_w, int n_b) {
  __shared__ float s_sum[16][65];
  const int o = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int idx = blockIdx.x * 64 + o;
  float acc = 0.0f;
  if (idx < n_w) {
    for (int k = g; k < runs; k += 16) acc += part[(size_t)k * n_w + idx];
  } else if (idx < n_w + n_b) {
    for (int k = g; k < runs; k += 16) acc += part[(size_t)runs * n_w + (size_t)k * n_b + (idx - n_w)];
  }
  s_sum[g][o] = acc;
  __syncthreads();
  if (g == 0) {
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += s_sum[k][o];
    if (idx < n_w) dw[idx] += t;
    else if (idx < n_w + n_b) db[idx - n_w] += t;
  }
}

__global__ void k_flip(const bf16* __restrict__ w, bf16* __restrict__ wt, int Cin, int Cout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = 9 * Cin * Cout;
  if (i >= n) return;
  // wt index: [r][s][c][k] with k fastest
  const int k = i % Cout, c = (i / Cout) % Cin, rs = i / (Cout * Cin);
  const int r = rs / 3, s = rs % 3;
  wt[i] = w[((size_t)((2 - r) * 3 + (2 - s)) * Cout + k) * Cin + c];
}

__global__ void k_relu_mask(View dy, View y, int npix, int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)npix * C) return;
  const size_t p = i / C, c = i % C;
  bf16* d = dy.p + p * dy.cs + dy.co + c;
  if (!(bf2f(y.p[p * y.cs + y.co + c]) > 0.0f)) *d = 0;
}

__global__ void k_avgpool2(View in, View out, int N, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * C) return;
  const int c = (int)(i % C);
  const size_t po = i / C;
  const int xo = (int)(po % Wo), yo = (int)((po / Wo) % Ho), n = (int)(po / ((size_t)Wo * Ho));
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) s += bf2f(in.p[((size_t)(n * H + 2 * yo + a) * W + 2 * xo + b) * in.cs + in.co + c]);
  out.p[po * out.cs + out.co + c] = f2bf(0.25f * s);
}

__global__ void k_avgpool2_bwd(View dout, View din, int N, int H, int W, int C, int acc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * H * W * C) return;
  const int c = (int)(i % C);
  const size_t p = i / C;
  const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
  const float g = 0.25f * bf2f(dout.p[((size_t)(n * (H / 2) + y / 2) * (W / 2) + x / 2) * dout.cs + dout.co + c]);
  bf16* d = din.p + p * din.cs + din.co + c;
  *d = f2bf(acc ? bf2f(*d) + g : g);
}

__global__ void k_upsample2(View in, View out, int N, int H, int W, int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = 2 * H, Wo = 2 * W;
  if (i >= (size_t)N * Ho * Wo * C) return;
  const int c = (int)(i % C);
  const size_t po = i / C;
  const int xo = (int)(po % Wo), yo = (int)((po / Wo) % Ho), n = (int)(po / ((size_t)Wo * Ho));
  out.p[po * out.cs + out.co + c] = in.p[((size_t)(n * H + yo / 2) * W + xo / 2) * in.cs + in.co + c];
}

// 8-channel (16-byte) versions of pooling and upsampling: one thread per
// output pixel and channel group.

__global__ void k_avgpool2_v8(View in, View out, int N, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2, G = C / 8;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * G) return;
  const int g = (int)(i % G);
  const size_t po = i / G;
  const int xo = (int)(po % Wo), yo = (int)((po / Wo) % Ho), n = (int)(po / ((size_t)Wo * Ho));
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const uint4 v = *reinterpret_cast<const uint4*>(
          in.p + ((size_t)(n * H + 2 * yo + a) * W + 2 * xo + b) * in.cs + in.co + 8 * g);
      const uint32_t q[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) { s[2 * k] += bf_lo(q[k]); s[2 * k + 1] += bf_hi(q[k]); }
    }
  uint32_t r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    r[k] = (uint32_t)f2bf(0.25f * s[2 * k]) | ((uint32_t)f2bf(0.25f * s[2 * k + 1]) << 16);
  *reinterpret_cast<uint4*>(out.p + po * out.cs + out.co + 8 * g) = make_uint4(r[0], r[1], r[2], r[3]);
}

__global__ void k_upsample2_v8(View in, View out, int N, int H, int W, int C) {
  const int Ho = 2 * H, Wo = 2 * W, G = C / 8;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * G) return;
  const int g = (int)(i % G);
  const size_t po = i / G;
  const int xo = (int)(po % Wo), yo = (int)((po / Wo) % Ho), n = (int)(po / ((size_t)Wo * Ho));
  *reinterpret_cast<uint4*>(out.p + po * out.cs + out.co + 8 * g) =
      *reinterpret_cast<const uint4*>(in.p + ((size_t)(n * H + yo / 2) * W + xo / 2) * in.cs + in.co + 8 * g);
}

inline bool v8_ok(const View& a, const View& b, int C) {
  return C % 8 == 0 && a.cs % 8 == 0 && a.co % 8 == 0 && b.cs % 8 == 0 && b.co % 8 == 0 &&
         ((uintptr_t)a.p & 15u) == 0 && ((uintptr_t)b.p & 15u) == 0;
}

__global__ void k_upsample2_bwd(View dout, View din, int N, int H, int W, int C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * H * W * C) return;
  const int c = (int)(i % C);
  const size_t p = i / C;
  const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
      s += bf2f(dout.p[((size_t)(n * 2 * H + 2 * y + a) * (2 * W) + 2 * x + b) * dout.cs + dout.co + c]);
  din.p[p * din.cs + din.co + c] = f2bf(s);
}

__global__ void k_residual_loss(View x, View o, const bf16* __restrict__ t, float* __restrict__ y_out,
                                float* __restrict__ loss_sum, bf16* __restrict__ dO, int npix, float inv_count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (p < npix) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float y = bf2f(x.p[(size_t)p * x.cs + x.co + c]) + bf2f(o.p[(size_t)p * o.cs + o.co + c]);
      if (y_out) y_out[(size_t)p * 4 + c] = y;
      if (t) {
        const float d = y - bf2f(t[(size_t)p * 16 + c]);
        l += d * d;
        dO[(size_t)p * 16 + c] = f2bf(2.0f * d * inv_count);
      }
    }
    if (t) {
#pragma unroll
      for (int c = 4; c < 16; ++c) dO[(size_t)p * 16 + c] = 0;
    }
  }
  if (t) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss_sum, l);
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16* __restrict__ pb, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  // bias-corrected step: c1 = 1 - b1^t, c2 = 1 - b2^t
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ in, bf16* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// Network input: channels 0-3 the normalised log-density log10(1 + v scale)
// of the bin's RGBA, channel 4 the noise level log10(1 + scale) (what one
// sample in a bin reads: scale = W H / n is the inverse samples per pixel),
// channels 5-15 zero.
__global__ void k_log_density(const float* __restrict__ hist, bf16* __restrict__ out, int npix, float scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * 16) return;
  const int p = i >> 4, c = i & 15;
  out[i] = c < 4 ? f2bf(log10f(1.0f + hist[(size_t)p * 4 + c] * scale))
                 : (c == 4 ? f2bf(log10f(1.0f + scale)) : (bf16)0);
}

// The same values, one pixel per thread: one 16-byte load of the bin and two
// 16-byte stores of its 16 channels (16-byte aligned buffers).
__global__ void k_log_density_v(const float4* __restrict__ hist, uint4* __restrict__ out, int npix, float scale) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const float4 h = hist[p];
  const uint32_t lo = (uint32_t)f2bf(log10f(1.0f + h.x * scale)) | ((uint32_t)f2bf(log10f(1.0f + h.y * scale)) << 16);
  const uint32_t hi = (uint32_t)f2bf(log10f(1.0f + h.z * scale)) | ((uint32_t)f2bf(log10f(1.0f + h.w * scale)) << 16);
  const uint32_t nl = (uint32_t)f2bf(log10f(1.0f + scale));
  out[2 * (size_t)p] = make_uint4(lo, hi, nl, 0u);
  out[2 * (size_t)p + 1] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_bf16_to_f32_4(const bf16* __restrict__ in, int cs, float* __restrict__ out, int npix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * 4) return;
  out[i] = bf2f(in[(size_t)(i >> 2) * cs + (i & 3)]);
}

__global__ void k_zero(float* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t conv3x3(View in, const bf16* w, const float* bias, View out, int N, int H, int W, int Cin, int Cout,
                   int relu, hipStream_t s, const ConvFuse& fuse) {
  const View& up = fuse.up;
  const int Cup = fuse.Cup;
  const View& pool = fuse.pool;
  if (Cin % 16 || Cout % 16 || in.cs % 4 || in.co % 4 || out.cs < out.co + Cout || in.cs < in.co + Cin - Cup)
    return hipErrorInvalidValue;
  const int nt = (Cout % 64 == 0) ? 4 : (Cout % 32 == 0) ? 2 : 1;
  const bool lds = Cin % 8 == 0 && in.cs % 8 == 0 && in.co % 8 == 0 && ((uintptr_t)in.p & 15u) == 0 &&
                   ((uintptr_t)w & 15u) == 0;
  if (Cup) {  // the fused upsampling exists on the LDS path only
    if (!lds || Cup < 0 || Cup % 32 || Cup >= Cin || H % 2 || W % 2 || up.cs % 8 || up.co % 8 ||
        up.cs < up.co + Cup || ((uintptr_t)up.p & 15u) != 0)
      return hipErrorInvalidValue;
  }
  if (lds) {
    // K = 32 steps (v_mfma_f32_16x16x32_bf16: twice the products per
    // instruction) pay off where each staged step feeds 64 output channels or
    // the input has 64 channels or more: the 96 -> 32 decoder layer at full
    // resolution takes 320 us at 2048^2 with them, 389 us without.  The
    // narrow layers with 16 or 32 input channels keep the smaller K = 16
    // tiles, whose LDS and registers fit three workgroups per CU instead of
    // two (K = 32 there measured 203 -> 205 and 167 -> 175 us for the 32 ->
    // 32 layers, 118 -> 144 us for the 32 -> 16 output layer).  The narrow
    // layers give each wave two rows (RW = 2), whose shared input rows halve
    // the B and cut the A fragment reads per output by a third.
    const bool k32 = Cin % 32 == 0 && (nt == 4 || Cin >= 64);
    const int rw = nt <= 2 ? 2 : 1;
    const int nbx = (W + CT_W - 1) / CT_W, nbz = Cout / (16 * nt);
    const long long tiles = (long long)nbx * N * ((H + 4 * rw - 1) / (4 * rw)) * nbz;
    if (tiles >= (1ll << 24)) return hipErrorInvalidValue;  // a 1-D grid of 256-thread workgroups
    const int n_tiles = (int)tiles;
    const int vec_out = out.cs % 8 == 0 && out.co % 8 == 0 && ((uintptr_t)out.p & 15u) == 0;
    // the fused pooling needs the LDS epilogue; otherwise pool afterwards
    const bool fuse_pool = pool.p && vec_out && H % 2 == 0 && W % 2 == 0 && pool.cs % 8 == 0 &&
                           pool.co % 8 == 0 && ((uintptr_t)pool.p & 15u) == 0;
    // the residual output needs the LDS epilogue of a single 16-channel block;
    // otherwise the output is written and the residual computed afterwards
    const bool fuse_res = fuse.y && vec_out && Cout == 16 && fuse.res.cs % 4 == 0 && fuse.res.co % 4 == 0 &&
                          ((uintptr_t)fuse.res.p & 7u) == 0 && ((uintptr_t)fuse.y & 15u) == 0 && !pool.p;
    ConvFuse fz = fuse;
    if (!fuse_pool) fz.pool = View{nullptr, 0, 0};
    if (!fuse_res) { fz.res = View{nullptr, 0, 0}; fz.y = nullptr; }
#define RF_CONV_LDS(NT_, KC_, RW_)                                                                            \
  hipLaunchKernelGGL((k_conv3x3_lds<NT_, KC_, RW_>), dim3((unsigned)n_tiles), dim3(256), 0, s, in, w, bias, out, H, \
                     W, Cin, Cout, relu, vec_out, fz, nbx, nbz, n_tiles)
    if (k32) {
      switch (nt) {
        case 4: RF_CONV_LDS(4, 32, 1); break;
        case 2: RF_CONV_LDS(2, 32, 2); break;
        default: RF_CONV_LDS(1, 32, 2); break;
      }
    } else {
      switch (nt) {
        case 4: RF_CONV_LDS(4, 16, 1); break;
        case 2: RF_CONV_LDS(2, 16, 2); break;
        default: RF_CONV_LDS(1, 16, 2); break;
      }
    }
#undef RF_CONV_LDS
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && pool.p && !fuse_pool) e = avgpool2(out, pool, N, H, W, Cout, s);
    if (e == hipSuccess && fuse.y && !fuse_res)
      e = residual_loss(fuse.res, out, nullptr, fuse.y, nullptr, nullptr, N * H * W, 0.0f, s);
    return e;
  }
  const dim3 grid((W + 63) / 64, N * H, Cout / (16 * nt));
  switch (nt) {
    case 4: hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(256), 0, s, in, w, bias, out, H, W, Cin, Cout, relu); break;
    case 2: hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, s, in, w, bias, out, H, W, Cin, Cout, relu); break;
    default: hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(256), 0, s, in, w, bias, out, H, W, Cin, Cout, relu); break;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && pool.p) e = avgpool2(out, pool, N, H, W, Cout, s);
  if (e == hipSuccess && fuse.y) e = residual_loss(fuse.res, out, nullptr, fuse.y, nullptr, nullptr, N * H * W, 0.0f, s);
  return e;
}

namespace {
// Channel tile of the weight gradient: TCO x TCI, SPW output blocks per wave.
struct WgradTile {
  int tco, tci;
};
WgradTile wgrad_tile(int Cin, int Cout) {
  const int tci = Cin % 32 == 0 ? 32 : 16;
  const int tco = (Cout % 64 == 0 && tci == 32) ? 64 : Cout % 32 == 0 ? 32 : 16;
  return WgradTile{tco, tci};
}
// Runs (row-strip ranges with their own partial sums): about 2048
// workgroups, at least 4 units per run, at most 2^23 partial sums.
int wgrad_runs(int N, int H, int W, int Cin, int Cout) {
  const WgradTile t = wgrad_tile(Cin, Cout);
  const int blocks = (Cout / t.tco) * (Cin / t.tci);
  const long long units = (long long)N * H * ((W + WG_SEG - 1) / WG_SEG);
  const size_t per = (size_t)9 * Cout * Cin + Cout;
  long long runs = std::min<long long>((2048 + blocks - 1) / blocks, std::max<long long>(1, units / 4));
  runs = std::min<long long>(runs, (long long)std::max<size_t>(1, ((size_t)1 << 23) / per));
  return (int)std::max<long long>(1, runs);
}

template <int TCO, int TCI, int SPW>
void launch_wgrad(dim3 grid, hipStream_t s, View x, View dy, float* work, int N, int H, int W, int Cin, int Cout,
                  int runs, int with_bias) {
  hipLaunchKernelGGL((k_wgrad_mfma<TCO, TCI, SPW>), grid, dim3(256), 0, s, x, dy, work, N, H, W, Cin, Cout, runs,
                     with_bias);
}
}  // namespace

size_t conv3x3_wgrad_work(int N, int H, int W, int Cin, int Cout) {
  return (size_t)wgrad_runs(N, H, W, Cin, Cout) * ((size_t)9 * Cout * Cin + Cout);
}

hipError_t conv3x3_wgrad(View x, View dy, float* dw, float* db, int N, int H, int W, int Cin, int Cout, float* work,
                         size_t work_floats, hipStream_t s) {
  if (Cin % 16 || Cout % 16 || N <= 0 || H <= 0 || W <= 0 || !work) return hipErrorInvalidValue;
  if (work_floats < conv3x3_wgrad_work(N, H, W, Cin, Cout)) return hipErrorInvalidValue;
  const WgradTile t = wgrad_tile(Cin, Cout);
  const int runs = wgrad_runs(N, H, W, Cin, Cout);
  const dim3 grid((unsigned)runs, (unsigned)(Cin / t.tci), (unsigned)(Cout / t.tco));
  const int wb = db ? 1 : 0;
  if (t.tco == 64) launch_wgrad<64, 32, 2>(grid, s, x, dy, work, N, H, W, Cin, Cout, runs, wb);
  else if (t.tco == 32 && t.tci == 32) launch_wgrad<32, 32, 1>(grid, s, x, dy, work, N, H, W, Cin, Cout, runs, wb);
  else if (t.tco == 32) launch_wgrad<32, 16, 1>(grid, s, x, dy, work, N, H, W, Cin, Cout, runs, wb);
  else if (t.tci == 32) launch_wgrad<16, 32, 1>(grid, s, x, dy, work, N, H, W, Cin, Cout, runs, wb);
  else launch_wgrad<16, 16, 1>(grid, s, x, dy, work, N, H, W, Cin, Cout, runs, wb);
  const int n_w = 9 * Cout * Cin, n_b = db ? Cout : 0;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((n_w + n_b + 63) / 64)), dim3(1024), 0, s, work, dw, db, runs,
                     n_w, n_b);
  return hipGetLastError();
}

hipError_t flip_weights(const bf16* w, bf16* wt, int Cin, int Cout, hipStream_t s) {
  hipLaunchKernelGGL(k_flip, grid1((size_t)9 * Cin * Cout), dim3(256), 0, s, w, wt, Cin, Cout);
  return hipGetLastError();
}

hipError_t relu_mask(View dy, View y, int npix, int C, hipStream_t s) {
  hipLaunchKernelGGL(k_relu_mask, grid1((size_t)npix * C), dim3(256), 0, s, dy, y, npix, C);
  return hipGetLastError();
}

hipError_t avgpool2(View in, View out, int N, int H, int W, int C, hipStream_t s) {
  if (v8_ok(in, out, C))
    hipLaunchKernelGGL(k_avgpool2_v8, grid1((size_t)N * (H / 2) * (W / 2) * (C / 8)), dim3(256), 0, s, in, out, N, H,
                       W, C);
  else
    hipLaunchKernelGGL(k_avgpool2, grid1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0, s, in, out, N, H, W, C);
  return hipGetLastError();
}

hipError_t avgpool2_bwd(View dout, View din, int N, int H, int W, int C, int acc, hipStream_t s) {
  hipLaunchKernelGGL(k_avgpool2_bwd, grid1((size_t)N * H * W * C), dim3(256), 0, s, dout, din, N, H, W, C, acc);
  return hipGetLastError();
}

hipError_t upsample2(View in, View out, int N, int H, int W, int C, hipStream_t s) {
  if (v8_ok(in, out, C))
    hipLaunchKernelGGL(k_upsample2_v8, grid1((size_t)N * 4 * H * W * (C / 8)), dim3(256), 0, s, in, out, N, H, W, C);
  else
    hipLaunchKernelGGL(k_upsample2, grid1((size_t)N * 4 * H * W * C), dim3(256), 0, s, in, out, N, H, W, C);
  return hipGetLastError();
}

hipError_t upsample2_bwd(View dout, View din, int N, int H, int W, int C, hipStream_t s) {
  hipLaunchKernelGGL(k_upsample2_bwd, grid1((size_t)N * H * W * C), dim3(256), 0, s, dout, din, N, H, W, C);
  return hipGetLastError();
}

hipError_t residual_loss(View x, View o, const bf16* t, float* y_out, float* loss_sum, bf16* dO, int npix,
                         float inv_count, hipStream_t s) {
  hipLaunchKernelGGL(k_residual_loss, grid1((size_t)npix), dim3(256), 0, s, x, o, t, y_out, loss_sum, dO, npix,
                     inv_count);
  return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m, float* v, bf16* pb, int n, float lr, float b1, float b2,
                float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, grid1((size_t)n), dim3(256), 0, s, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t to_bf16(const float* in, bf16* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, grid1((size_t)n), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t log_density_bf16(const float* hist, bf16* out, int npix, float scale, hipStream_t s) {
  if (npix < 0 || (size_t)npix * 16 > (size_t)INT32_MAX) return hipErrorInvalidValue;
  if ((((uintptr_t)hist | (uintptr_t)out) & 15u) == 0)
    hipLaunchKernelGGL(k_log_density_v, grid1((size_t)npix), dim3(256), 0, s, reinterpret_cast<const float4*>(hist),
                       reinterpret_cast<uint4*>(out), npix, scale);
  else
    hipLaunchKernelGGL(k_log_density, grid1((size_t)npix * 16), dim3(256), 0, s, hist, out, npix, scale);
  return hipGetLastError();
}

hipError_t bf16_to_f32_4(const bf16* in, int cs, float* out, int npix, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_to_f32_4, grid1((size_t)npix * 4), dim3(256), 0, s, in, cs, out, npix);
  return hipGetLastError();
}

hipError_t zero_f32(float* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, grid1(n), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace rfn
