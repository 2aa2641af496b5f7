// Learned denoiser behind the C ABI: a 4-level U-Net on the normalised
// log-density image, bf16 activations, fp32 master weights, Adam.
//
// Network (3x3 convolutions, ReLU after every one but the last; channels):
//   enc0: 16 -> 32 -> 32 (skip0)    pool  enc1: 32 -> 64 -> 64 (skip1)
//   pool  enc2: 64 -> 128 -> 128 (skip2)  pool  bottleneck: 128 -> 256
//   up + skip2: 384 -> 128   up + skip1: 192 -> 64   up + skip0: 96 -> 32 -> 32
//   out: 32 -> 16, and the result is input[:4] + out[:4]
// The input holds the 4 channels of log10(1 + v * W * H / n) of a histogram
// (zero-padded to 16 channels); 1.18 M parameters.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rf_host.hpp"
#include "rf_nn.hpp"

using rfh::DevBuf;
using rfn::bf16;
using rfn::View;

namespace {

struct Layer {
  int cin, cout, relu;
  size_t w_off, b_off;  // into the flat parameter vector
};

const int kSpec[][3] = {
    {16, 32, 1}, {32, 32, 1},     // enc0
    {32, 64, 1}, {64, 64, 1},     // enc1
    {64, 128, 1}, {128, 128, 1},  // enc2
    {128, 256, 1},                // bottleneck
    {384, 128, 1},                // dec2
    {192, 64, 1},                 // dec1
    {96, 32, 1}, {32, 32, 1},     // dec0
    {32, 16, 0},                  // out
};
constexpr int kLayers = 12;

}  // namespace

struct rf_denoiser {
  rf_ctx* ctx = nullptr;
  std::vector<Layer> L;
  size_t n_params = 0;
  DevBuf p, g, m, v, pb, wt;   // fp32 params/grads/moments, bf16 params, bf16 flipped weights
  std::vector<size_t> wt_off;  // per layer, into wt
  DevBuf loss;
  int step = 0;
  // activations and gradients for the current shape
  int N = 0, H = 0, W = 0;
  DevBuf e0a, cat0, p0, e1a, cat1, p1, e2a, cat2, p2, b, d2, d1, d0a, d0b, o;
  DevBuf g_e0a, g_cat0, g_p0, g_e1a, g_cat1, g_p1, g_e2a, g_cat2, g_p2, g_b, g_d2, g_d1, g_d0a, g_d0b, g_o;
  DevBuf wg_work;  // per-split weight-gradient sums of one layer (fp32)
};

namespace {

int ensure_shape(rf_denoiser* d, int N, int H, int W) {
  rf_ctx* ctx = d->ctx;
  if (N == d->N && H == d->H && W == d->W) return RF_OK;
  const size_t P0 = (size_t)N * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
  struct A { DevBuf* b; size_t n; };
  const A all[] = {
      {&d->e0a, P0 * 32}, {&d->cat0, P0 * 96}, {&d->p0, P1 * 32}, {&d->e1a, P1 * 64}, {&d->cat1, P1 * 192},
      {&d->p1, P2 * 64}, {&d->e2a, P2 * 128}, {&d->cat2, P2 * 384}, {&d->p2, P3 * 128}, {&d->b, P3 * 256},
      {&d->d2, P2 * 128}, {&d->d1, P1 * 64}, {&d->d0a, P0 * 32}, {&d->d0b, P0 * 32}, {&d->o, P0 * 16},
      {&d->g_e0a, P0 * 32}, {&d->g_cat0, P0 * 96}, {&d->g_p0, P1 * 32}, {&d->g_e1a, P1 * 64},
      {&d->g_cat1, P1 * 192}, {&d->g_p1, P2 * 64}, {&d->g_e2a, P2 * 128}, {&d->g_cat2, P2 * 384},
      {&d->g_p2, P3 * 128}, {&d->g_b, P3 * 256}, {&d->g_d2, P2 * 128}, {&d->g_d1, P1 * 64},
      {&d->g_d0a, P0 * 32}, {&d->g_d0b, P0 * 32}, {&d->g_o, P0 * 16},
  };
  for (const A& a : all) HIPCHK(ctx, a.b->alloc(a.n * sizeof(bf16)));
  // layer l runs at 1/2^level of the image height
  const int level[kLayers] = {0, 0, 1, 1, 2, 2, 3, 2, 1, 0, 0, 0};
  size_t work = 0;
  for (int l = 0; l < kLayers; ++l)
    work = std::max(work, rfn::conv3x3_wgrad_work(N, H >> level[l], W >> level[l], d->L[l].cin, d->L[l].cout));
  HIPCHK(ctx, d->wg_work.alloc(work * sizeof(float)));
  d->N = N; d->H = H; d->W = W;
  return RF_OK;
}

inline View V(DevBuf& b, int cs, int co = 0) { return View{(bf16*)b.p, cs, co}; }
inline const float* bias_of(rf_denoiser* d, int l) { return (const float*)d->p.p + d->L[l].b_off; }
inline const bf16* w_of(rf_denoiser* d, int l) { return (const bf16*)d->pb.p + d->L[l].w_off; }
inline const bf16* wt_of(rf_denoiser* d, int l) { return (const bf16*)d->wt.p + d->wt_off[l]; }

int refresh_weights(rf_denoiser* d) {
  rf_ctx* ctx = d->ctx;
  hipStream_t s = ctx->stream;
  HIPCHK(ctx, rfn::to_bf16((const float*)d->p.p, (bf16*)d->pb.p, (int)d->n_params, s));
  for (int l = 0; l < kLayers; ++l)
    HIPCHK(ctx, rfn::flip_weights(w_of(d, l), (bf16*)wt_of(d, l), d->L[l].cin, d->L[l].cout, s));
  return RF_OK;
}

int check_shape(rf_denoiser* d, int N, int H, int W) {
  if (N <= 0 || H <= 0 || W <= 0 || H % 8 || W % 8)
    return rfh::set_err(d->ctx, RF_ERR_ARG, "denoiser: N > 0 and H, W multiples of 8 required");
  if ((double)N * H * W * 96 * 2 > 64.0 * (1ull << 30))
    return rfh::set_err(d->ctx, RF_ERR_LIMIT, "denoiser: batch too large for the activation buffers");
  return RF_OK;
}

// Forward pass on x (N x H x W x 16 bf16).  With y_out (inference) the
// decoder's convolutions read the upsampled half of their concatenated input
// straight from the lower level and the last one writes y_out = x[:4] +
// out[:4] (fp32); training writes the upsampled halves into the
// concatenation buffers, whose full contents the weight gradients read back,
// and keeps the last layer's output for the loss.
int forward(rf_denoiser* d, const bf16* x, int N, int H, int W, float* y_out = nullptr) {
  const bool fused_up = y_out != nullptr;
  rf_ctx* ctx = d->ctx;
  hipStream_t s = ctx->stream;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  const View vx{(bf16*)x, 16, 0};
  auto conv = [&](int l, View in, View out, int h, int w) {
    return rfn::conv3x3(in, w_of(d, l), bias_of(d, l), out, N, h, w, d->L[l].cin, d->L[l].cout, d->L[l].relu, s);
  };
  // conv l that also writes the 2x2 average of its output to `pool`
  auto conv_f = [&](int l, View in, View out, int h, int w, const rfn::ConvFuse& f) {
    return rfn::conv3x3(in, w_of(d, l), bias_of(d, l), out, N, h, w, d->L[l].cin, d->L[l].cout, d->L[l].relu, s, f);
  };
  auto conv_pool = [&](int l, View in, View out, View pool, int h, int w) {
    rfn::ConvFuse f;
    f.pool = pool;
    return conv_f(l, in, out, h, w, f);
  };
  // The skip channels of a concatenation buffer `cat` (cat_cs channels, the
  // upsampled c_lo first): inside it for training, packed at their own
  // channel count at inference, where no upsampled half sits beside them
  // (whole-line writes and reads instead of a third or half of each pixel).
  auto skip = [&](DevBuf& cat, int cat_cs, int c_lo) {
    return fused_up ? V(cat, cat_cs - c_lo, 0) : V(cat, cat_cs, c_lo);
  };
  // conv l over [upsample(lo) (c_lo channels), the skip channels of `cat`]
  auto up_conv = [&](int l, View lo, int c_lo, DevBuf& cat, int cat_cs, View out, int h, int w) -> hipError_t {
    if (fused_up) {
      rfn::ConvFuse f;
      f.up = lo;
      f.Cup = c_lo;
      return conv_f(l, skip(cat, cat_cs, c_lo), out, h, w, f);
    }
    const hipError_t e = rfn::upsample2(lo, V(cat, cat_cs, 0), N, h / 2, w / 2, c_lo, s);
    return e != hipSuccess ? e : conv(l, V(cat, cat_cs), out, h, w);
  };
  HIPCHK(ctx, conv(0, vx, V(d->e0a, 32), H, W));
  HIPCHK(ctx, conv_pool(1, V(d->e0a, 32), skip(d->cat0, 96, 64), V(d->p0, 32), H, W));
  HIPCHK(ctx, conv(2, V(d->p0, 32), V(d->e1a, 64), H1, W1));
  HIPCHK(ctx, conv_pool(3, V(d->e1a, 64), skip(d->cat1, 192, 128), V(d->p1, 64), H1, W1));
  HIPCHK(ctx, conv(4, V(d->p1, 64), V(d->e2a, 128), H2, W2));
  HIPCHK(ctx, conv_pool(5, V(d->e2a, 128), skip(d->cat2, 384, 256), V(d->p2, 128), H2, W2));
  HIPCHK(ctx, conv(6, V(d->p2, 128), V(d->b, 256), H3, W3));
  HIPCHK(ctx, up_conv(7, V(d->b, 256), 256, d->cat2, 384, V(d->d2, 128), H2, W2));
  HIPCHK(ctx, up_conv(8, V(d->d2, 128), 128, d->cat1, 192, V(d->d1, 64), H1, W1));
  HIPCHK(ctx, up_conv(9, V(d->d1, 64), 64, d->cat0, 96, V(d->d0a, 32), H, W));
  HIPCHK(ctx, conv(10, V(d->d0a, 32), V(d->d0b, 32), H, W));
  if (y_out) {
    rfn::ConvFuse f;
    f.res = vx;
    f.y = y_out;
    HIPCHK(ctx, conv_f(11, V(d->d0b, 32), V(d->o, 16), H, W, f));
  } else {
    HIPCHK(ctx, conv(11, V(d->d0b, 32), V(d->o, 16), H, W));
  }
  return RF_OK;
}

// Backward pass from g_o (dLoss/dout, 16 channels); gradients into d->g.
int backward(rf_denoiser* d, const bf16* x, int N, int H, int W) {
  rf_ctx* ctx = d->ctx;
  hipStream_t s = ctx->stream;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  const size_t P0 = (size_t)N * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
  float* g = (float*)d->g.p;
  HIPCHK(ctx, rfn::zero_f32(g, d->n_params, s));
  auto wgrad = [&](int l, View xin, View dy, int h, int w) {
    return rfn::conv3x3_wgrad(xin, dy, g + d->L[l].w_off, g + d->L[l].b_off, N, h, w, d->L[l].cin, d->L[l].cout,
                              (float*)d->wg_work.p, d->wg_work.bytes / sizeof(float), s);
  };
  auto dgrad = [&](int l, View dy, View dx, int h, int w) {
    return rfn::conv3x3(dy, wt_of(d, l), nullptr, dx, N, h, w, d->L[l].cout, d->L[l].cin, 0, s);
  };
  const View vx{(bf16*)x, 16, 0};
  // out and dec0
  HIPCHK(ctx, wgrad(11, V(d->d0b, 32), V(d->g_o, 16), H, W));
  HIPCHK(ctx, dgrad(11, V(d->g_o, 16), V(d->g_d0b, 32), H, W));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_d0b, 32), V(d->d0b, 32), (int)P0, 32, s));
  HIPCHK(ctx, wgrad(10, V(d->d0a, 32), V(d->g_d0b, 32), H, W));
  HIPCHK(ctx, dgrad(10, V(d->g_d0b, 32), V(d->g_d0a, 32), H, W));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_d0a, 32), V(d->d0a, 32), (int)P0, 32, s));
  HIPCHK(ctx, wgrad(9, V(d->cat0, 96), V(d->g_d0a, 32), H, W));
  HIPCHK(ctx, dgrad(9, V(d->g_d0a, 32), V(d->g_cat0, 96), H, W));
  // dec1
  HIPCHK(ctx, rfn::upsample2_bwd(V(d->g_cat0, 96, 0), V(d->g_d1, 64), N, H1, W1, 64, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_d1, 64), V(d->d1, 64), (int)P1, 64, s));
  HIPCHK(ctx, wgrad(8, V(d->cat1, 192), V(d->g_d1, 64), H1, W1));
  HIPCHK(ctx, dgrad(8, V(d->g_d1, 64), V(d->g_cat1, 192), H1, W1));
  // dec2
  HIPCHK(ctx, rfn::upsample2_bwd(V(d->g_cat1, 192, 0), V(d->g_d2, 128), N, H2, W2, 128, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_d2, 128), V(d->d2, 128), (int)P2, 128, s));
  HIPCHK(ctx, wgrad(7, V(d->cat2, 384), V(d->g_d2, 128), H2, W2));
  HIPCHK(ctx, dgrad(7, V(d->g_d2, 128), V(d->g_cat2, 384), H2, W2));
  // bottleneck
  HIPCHK(ctx, rfn::upsample2_bwd(V(d->g_cat2, 384, 0), V(d->g_b, 256), N, H3, W3, 256, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_b, 256), V(d->b, 256), (int)P3, 256, s));
  HIPCHK(ctx, wgrad(6, V(d->p2, 128), V(d->g_b, 256), H3, W3));
  HIPCHK(ctx, dgrad(6, V(d->g_b, 256), V(d->g_p2, 128), H3, W3));
  // enc2: skip2 gradient = concat part + pooled path
  HIPCHK(ctx, rfn::avgpool2_bwd(V(d->g_p2, 128), V(d->g_cat2, 384, 256), N, H2, W2, 128, 1, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_cat2, 384, 256), V(d->cat2, 384, 256), (int)P2, 128, s));
  HIPCHK(ctx, wgrad(5, V(d->e2a, 128), V(d->g_cat2, 384, 256), H2, W2));
  HIPCHK(ctx, dgrad(5, V(d->g_cat2, 384, 256), V(d->g_e2a, 128), H2, W2));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_e2a, 128), V(d->e2a, 128), (int)P2, 128, s));
  HIPCHK(ctx, wgrad(4, V(d->p1, 64), V(d->g_e2a, 128), H2, W2));
  HIPCHK(ctx, dgrad(4, V(d->g_e2a, 128), V(d->g_p1, 64), H2, W2));
  // enc1
  HIPCHK(ctx, rfn::avgpool2_bwd(V(d->g_p1, 64), V(d->g_cat1, 192, 128), N, H1, W1, 64, 1, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_cat1, 192, 128), V(d->cat1, 192, 128), (int)P1, 64, s));
  HIPCHK(ctx, wgrad(3, V(d->e1a, 64), V(d->g_cat1, 192, 128), H1, W1));
  HIPCHK(ctx, dgrad(3, V(d->g_cat1, 192, 128), V(d->g_e1a, 64), H1, W1));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_e1a, 64), V(d->e1a, 64), (int)P1, 64, s));
  HIPCHK(ctx, wgrad(2, V(d->p0, 32), V(d->g_e1a, 64), H1, W1));
  HIPCHK(ctx, dgrad(2, V(d->g_e1a, 64), V(d->g_p0, 32), H1, W1));
  // enc0
  HIPCHK(ctx, rfn::avgpool2_bwd(V(d->g_p0, 32), V(d->g_cat0, 96, 64), N, H, W, 32, 1, s));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_cat0, 96, 64), V(d->cat0, 96, 64), (int)P0, 32, s));
  HIPCHK(ctx, wgrad(1, V(d->e0a, 32), V(d->g_cat0, 96, 64), H, W));
  HIPCHK(ctx, dgrad(1, V(d->g_cat0, 96, 64), V(d->g_e0a, 32), H, W));
  HIPCHK(ctx, rfn::relu_mask(V(d->g_e0a, 32), V(d->e0a, 32), (int)P0, 32, s));
  HIPCHK(ctx, wgrad(0, vx, V(d->g_e0a, 32), H, W));
  return RF_OK;
}

}  // namespace

extern "C" {

int rf_denoiser_create(rf_ctx* ctx, uint64_t seed, rf_denoiser** out) {
  if (!ctx || !out) return RF_ERR_ARG;
  *out = nullptr;
  rf_denoiser* d = new rf_denoiser();
  d->ctx = ctx;
  size_t off = 0, woff = 0;
  for (int l = 0; l < kLayers; ++l) {
    Layer y{kSpec[l][0], kSpec[l][1], kSpec[l][2], 0, 0};
    y.w_off = off;
    off += (size_t)9 * y.cin * y.cout;
    y.b_off = off;
    off += (size_t)y.cout;
    d->L.push_back(y);
    d->wt_off.push_back(woff);
    woff += (size_t)9 * y.cin * y.cout;
  }
  d->n_params = off;
  // He initialisation; the last layer starts small so the net starts near identity.
  std::vector<float> h(off, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    const Layer& y = d->L[l];
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * (float)y.cin)) * (l == kLayers - 1 ? 0.1f : 1.0f));
    for (size_t i = 0; i < (size_t)9 * y.cin * y.cout; ++i) h[y.w_off + i] = nd(rng);
  }
  hipError_t e = d->p.alloc(off * 4);
  if (e == hipSuccess) e = d->g.alloc(off * 4);
  if (e == hipSuccess) e = d->m.alloc(off * 4);
  if (e == hipSuccess) e = d->v.alloc(off * 4);
  if (e == hipSuccess) e = d->pb.alloc(off * 2);
  if (e == hipSuccess) e = d->wt.alloc(woff * 2);
  if (e == hipSuccess) e = d->loss.alloc(4);
  if (e == hipSuccess) e = hipMemcpy(d->p.p, h.data(), off * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(d->m.p, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(d->v.p, 0, off * 4);
  if (e != hipSuccess) {
    delete d;
    return rfh::set_err(ctx, RF_ERR_HIP, std::string("denoiser: ") + hipGetErrorString(e));
  }
  const int rc = refresh_weights(d);
  if (rc) { delete d; return rc; }
  *out = d;
  return RF_OK;
}

void rf_denoiser_destroy(rf_denoiser* d) { delete d; }

size_t rf_denoiser_num_params(const rf_denoiser* d) { return d ? d->n_params : 0; }

int rf_denoiser_get_params(rf_denoiser* d, float* host, size_t n) {
  if (!d || !host) return RF_ERR_ARG;
  if (n != d->n_params) return rfh::set_err(d->ctx, RF_ERR_ARG, "denoiser: parameter count mismatch");
  HIPCHK(d->ctx, hipStreamSynchronize(d->ctx->stream));
  HIPCHK(d->ctx, hipMemcpy(host, d->p.p, n * 4, hipMemcpyDeviceToHost));
  return RF_OK;
}

int rf_denoiser_set_params(rf_denoiser* d, const float* host, size_t n) {
  if (!d || !host) return RF_ERR_ARG;
  if (n != d->n_params) return rfh::set_err(d->ctx, RF_ERR_ARG, "denoiser: parameter count mismatch");
  HIPCHK(d->ctx, hipStreamSynchronize(d->ctx->stream));
  HIPCHK(d->ctx, hipMemcpy(d->p.p, host, n * 4, hipMemcpyHostToDevice));
  HIPCHK(d->ctx, hipMemset(d->m.p, 0, n * 4));
  HIPCHK(d->ctx, hipMemset(d->v.p, 0, n * 4));
  d->step = 0;
  return refresh_weights(d);
}

int rf_denoiser_train_step(rf_denoiser* d, const void* x, const void* target, int32_t N, int32_t H, int32_t W,
                           float lr, float* loss_out) {
  if (!d || !x || !target) return RF_ERR_ARG;
  int rc = check_shape(d, N, H, W);
  if (rc) return rc;
  if ((rc = ensure_shape(d, N, H, W))) return rc;
  rf_ctx* ctx = d->ctx;
  hipStream_t s = ctx->stream;
  if ((rc = forward(d, (const bf16*)x, N, H, W))) return rc;
  const int npix = N * H * W;
  HIPCHK(ctx, rfn::zero_f32((float*)d->loss.p, 1, s));
  HIPCHK(ctx, rfn::residual_loss(View{(bf16*)x, 16, 0}, V(d->o, 16), (const bf16*)target, nullptr,
                                 (float*)d->loss.p, (bf16*)d->g_o.p, npix, 1.0f / (4.0f * (float)npix), s));
  if ((rc = backward(d, (const bf16*)x, N, H, W))) return rc;
  d->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)d->step), c2 = 1.0f - std::pow(b2, (float)d->step);
  HIPCHK(ctx, rfn::adam((float*)d->p.p, (const float*)d->g.p, (float*)d->m.p, (float*)d->v.p, (bf16*)d->pb.p,
                        (int)d->n_params, lr, b1, b2, 1e-8f, c1, c2, s));
  for (int l = 0; l < kLayers; ++l)
    HIPCHK(ctx, rfn::flip_weights(w_of(d, l), (bf16*)wt_of(d, l), d->L[l].cin, d->L[l].cout, s));
  if (loss_out) {
    float ls = 0.0f;
    HIPCHK(ctx, hipMemcpyAsync(&ls, d->loss.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    *loss_out = ls / (4.0f * (float)npix);
  }
  return RF_OK;
}

int rf_denoiser_infer(rf_denoiser* d, const void* x, float* y, int32_t N, int32_t H, int32_t W) {
  if (!d || !x || !y) return RF_ERR_ARG;
  int rc = check_shape(d, N, H, W);
  if (rc) return rc;
  if ((rc = ensure_shape(d, N, H, W))) return rc;
  if ((rc = forward(d, (const bf16*)x, N, H, W, y))) return rc;
  HIPCHK(d->ctx, hipStreamSynchronize(d->ctx->stream));
  return RF_OK;
}

int rf_denoiser_grads(rf_denoiser* d, const void* x, const void* target, int32_t N, int32_t H, int32_t W,
                      float* grads_host, float* loss_out) {
  if (!d || !x || !target || !grads_host) return RF_ERR_ARG;
  int rc = check_shape(d, N, H, W);
  if (rc) return rc;
  if ((rc = ensure_shape(d, N, H, W))) return rc;
  rf_ctx* ctx = d->ctx;
  hipStream_t s = ctx->stream;
  if ((rc = forward(d, (const bf16*)x, N, H, W))) return rc;
  const int npix = N * H * W;
  HIPCHK(ctx, rfn::zero_f32((float*)d->loss.p, 1, s));
  HIPCHK(ctx, rfn::residual_loss(View{(bf16*)x, 16, 0}, V(d->o, 16), (const bf16*)target, nullptr,
                                 (float*)d->loss.p, (bf16*)d->g_o.p, npix, 1.0f / (4.0f * (float)npix), s));
  if ((rc = backward(d, (const bf16*)x, N, H, W))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(s));
  HIPCHK(ctx, hipMemcpy(grads_host, d->g.p, d->n_params * 4, hipMemcpyDeviceToHost));
  if (loss_out) {
    float ls = 0.0f;
    HIPCHK(ctx, hipMemcpy(&ls, d->loss.p, 4, hipMemcpyDeviceToHost));
    *loss_out = ls / (4.0f * (float)npix);
  }
  return RF_OK;
}

int rf_log_density_input(rf_ctx* ctx, const void* hist, void* out, uint32_t npix, double scale) {
  if (!ctx || !hist || !out) return RF_ERR_ARG;
  HIPCHK(ctx, rfn::log_density_bf16((const float*)hist, (bf16*)out, (int)npix, (float)scale, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

// ------------------------------------------------ kernel-level diagnostics
int rf_nn_conv3x3(rf_ctx* ctx, const void* in, int32_t in_cs, int32_t in_co, const void* w, const float* bias,
                  void* out, int32_t out_cs, int32_t out_co, int32_t N, int32_t H, int32_t W, int32_t Cin,
                  int32_t Cout, int32_t relu) {
  if (!ctx || !in || !w || !out) return RF_ERR_ARG;
  HIPCHK(ctx, rfn::conv3x3(View{(bf16*)in, in_cs, in_co}, (const bf16*)w, bias, View{(bf16*)out, out_cs, out_co},
                           N, H, W, Cin, Cout, relu, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_nn_conv3x3_wgrad(rf_ctx* ctx, const void* x, int32_t x_cs, int32_t x_co, const void* dy, int32_t dy_cs,
                        int32_t dy_co, float* dw, float* db, int32_t N, int32_t H, int32_t W, int32_t Cin,
                        int32_t Cout) {
  if (!ctx || !x || !dy || !dw) return RF_ERR_ARG;
  if (N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || Cin % 16 || Cout % 16) return RF_ERR_ARG;
  DevBuf work;
  const size_t n_work = rfn::conv3x3_wgrad_work(N, H, W, Cin, Cout);
  HIPCHK(ctx, work.alloc(n_work * sizeof(float)));
  HIPCHK(ctx, rfn::conv3x3_wgrad(View{(bf16*)x, x_cs, x_co}, View{(bf16*)dy, dy_cs, dy_co}, dw, db, N, H, W, Cin,
                                 Cout, (float*)work.p, n_work, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_nn_flip_weights(rf_ctx* ctx, const void* w, void* wt, int32_t Cin, int32_t Cout) {
  if (!ctx || !w || !wt) return RF_ERR_ARG;
  HIPCHK(ctx, rfn::flip_weights((const bf16*)w, (bf16*)wt, Cin, Cout, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_nn_adam(rf_ctx* ctx, float* p, const float* g, float* m, float* v, int32_t n, float lr, float b1, float b2,
               float eps, int32_t step) {
  if (!ctx || !p || !g || !m || !v || step < 1) return RF_ERR_ARG;
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  HIPCHK(ctx, rfn::adam(p, g, m, v, nullptr, n, lr, b1, b2, eps, c1, c2, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

}  // extern "C"
