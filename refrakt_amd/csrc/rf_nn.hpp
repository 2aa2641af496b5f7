// Denoiser network kernels: interface between rf_denoiser.cpp and rf_k_nn.hip.
//
// Activations are NHWC bf16.  A "view" (ptr, cs, co) addresses channels
// [co, co + C) of a tensor whose pixels are cs channels apart, so skip
// connections are written straight into the channel range of the
// concatenated tensor that consumes them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rfn {

typedef uint16_t bf16;

struct View {
  bf16* p;
  int cs;   // channel stride (channels per pixel of the underlying tensor)
  int co;   // first channel of the view
};

// 3x3 convolution, stride 1, zero padding 1:
//   out[n,y,x,k] = act(bias[k] + sum_{r,s,c} in[n, y+r-1, x+s-1, c] * w[r][s][k][c])
// w: bf16 [3][3][Cout][Cin]; bias: fp32 [Cout] or null; act: ReLU if relu.
// Cin and Cout multiples of 16, channel strides/offsets multiples of 4.
struct ConvFuse {
  // Cup > 0: the first Cup input channels are the nearest-neighbour 2x
  // upsampling of `up` (an N x H/2 x W/2 tensor), read in place, and `in`
  // holds channels Cup..Cin-1 (Cup a multiple of 32, H and W even, 16-byte
  // aligned channel runs in both views).
  View up{nullptr, 0, 0};
  int Cup = 0;
  // pool.p set: the 2x2 average of the output (as avgpool2 computes it) is
  // also written to `pool`.
  View pool{nullptr, 0, 0};
  // y set: instead of `out`, y[p][c] = res[p][c] + out[p][c] for c < 4
  // (fp32, 4 channels per pixel, the output rounded to bf16 first, as
  // residual_loss computes it).
  View res{nullptr, 0, 0};
  float* y = nullptr;
};
hipError_t conv3x3(View in, const bf16* w, const float* bias, View out, int N, int H, int W, int Cin, int Cout,
                   int relu, hipStream_t s, const ConvFuse& fuse = ConvFuse{});
// Weight gradient: dw[r][s][k][c] += sum_p dy[p,k] * x[p + (r-1, s-1), c] and
// db[k] += sum_p dy[p,k] (bf16 products on the matrix cores, fp32 sums).  The
// image rows are split over workgroups whose partial sums go to `work`
// (conv3x3_wgrad_work(N, H, W, Cin, Cout) floats), then are added to dw/db in a
// fixed order: deterministic, no atomics.
size_t conv3x3_wgrad_work(int N, int H, int W, int Cin, int Cout);
hipError_t conv3x3_wgrad(View x, View dy, float* dw, float* db, int N, int H, int W, int Cin, int Cout, float* work,
                         size_t work_floats, hipStream_t s);
// w_t[r][s][c][k] = w[2-r][2-s][k][c]: the weights of the input-gradient
// convolution (dx = conv3x3(dy, w_t) with Cin' = Cout, Cout' = Cin).
hipError_t flip_weights(const bf16* w, bf16* wt, int Cin, int Cout, hipStream_t s);
// dy *= (y > 0) on a C-channel view pair.
hipError_t relu_mask(View dy, View y, int npix, int C, hipStream_t s);
// 2x2 average pooling and its gradient (accumulated into dx when acc).
hipError_t avgpool2(View in, View out, int N, int H, int W, int C, hipStream_t s);
hipError_t avgpool2_bwd(View dout, View din, int N, int H, int W, int C, int acc, hipStream_t s);
// Nearest 2x upsampling (H, W are the input size) and its gradient.
hipError_t upsample2(View in, View out, int N, int H, int W, int C, hipStream_t s);
hipError_t upsample2_bwd(View dout, View din, int N, int H, int W, int C, hipStream_t s);
// y = x[:, :4] + o[:, :4] (fp32, 4 channels, when y_out != null); when t != null: loss_sum += sum (y - t)^2 and
// do = 2 (y - t) / count in channels 0..3, 0 in the rest (16-channel do).
hipError_t residual_loss(View x, View o, const bf16* t, float* y_out, float* loss_sum, bf16* dO, int npix,
                         float inv_count, hipStream_t s);
// Adam step on fp32 parameters; writes the bf16 copy used by the kernels.
hipError_t adam(float* p, const float* g, float* m, float* v, bf16* p_bf16, int n, float lr, float b1, float b2,
                float eps, float c1, float c2, hipStream_t s);
hipError_t to_bf16(const float* in, bf16* out, int n, hipStream_t s);
// Normalised log-density input: out[p, c] = log10(1 + hist[p, c] * scale)
// for c < 4 (bins layout), 0 for c in 4..15; 16-channel bf16.
hipError_t log_density_bf16(const float* hist, bf16* out, int npix, float scale, hipStream_t s);
// out[p, c] = float(in[p, c]) for c < 4 (4-channel fp32).
hipError_t bf16_to_f32_4(const bf16* in, int cs, float* out, int npix, hipStream_t s);
hipError_t zero_f32(float* p, size_t n, hipStream_t s);

}  // namespace rfn
