// Batched rendering behind the C ABI: frame table, graph capture, autotuner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rf_batch.hpp"
#include "rf_host.hpp"
#include "rf_internal.hpp"
#include "rf_jit.hpp"

using rfd::Block;
using rfh::DevBuf;

struct rf_batch {
  rf_ctx* ctx = nullptr;
  uint32_t n = 0;
  std::vector<rf_flame_desc> desc;
  std::vector<uint32_t> W, H;
  std::vector<uint64_t> samples;
  std::vector<size_t> off;         // first bin of each frame in hist/images
  size_t total_bins = 0;
  DevBuf blocks, palettes, frames, hist, images, perms;
  DevBuf drawn, remaining, done_ts, items, tiles_done, de_done_ts, t0;
  DevBuf de_rad, de_wt;            // density estimation: radius per bin (largest frame), weight table
  int perm_waves = 0;
  rf_batch_config cfg{8, RF_HIST_LDS, 256, 20};
  double sample_scale = 1.0;       // < 1 while the autotuner times a reduced copy
  uint32_t n_items = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double clock_khz = 0.0;
  uint32_t launch = 0;             // RNG stream key of this batch (same for every shape)
  bool lds_ok = true;              // every frame allows the fixed-point LDS path
  bool ran = false;
  // Flame program: the frame with the most xforms, when its structure covers
  // every frame's (xform i runs the same variations in every frame).
  std::vector<Block> prog_block;
  int program_used = -1;
  double program_compile_s = 0.0;

  ~rf_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) { (void)hipStreamSynchronize(stream); (void)hipStreamDestroy(stream); }
  }
};

namespace {

constexpr uint32_t kMaxFrames = 4096;

int check_cfg(rf_ctx* ctx, const rf_batch_config& c) {
  if (c.block_waves != 4 && c.block_waves != 8) return rfh::set_err(ctx, RF_ERR_ARG, "batch: block_waves must be 4 or 8");
  if (c.hist_mode != RF_HIST_QUAD && c.hist_mode != RF_HIST_LDS)
    return rfh::set_err(ctx, RF_ERR_ARG, "batch: hist_mode must be RF_HIST_QUAD or RF_HIST_LDS");
  if (c.iters == 0 || c.iters > (1u << 20)) return rfh::set_err(ctx, RF_ERR_ARG, "batch: iters must be in 1..2^20");
  if (c.fuse > 1024) return rfh::set_err(ctx, RF_ERR_ARG, "batch: fuse must be at most 1024");
  return RF_OK;
}

// Lays out the work items for the current config and sample scale, uploads
// the frame table and captures the graph.
int capture(rf_batch* b) {
  rf_ctx* ctx = b->ctx;
  const rf_batch_config& c = b->cfg;
  const uint64_t N = 64ull * (uint64_t)c.block_waves;
  std::vector<rfk::BatchFrame> ft(b->n);
  std::vector<uint32_t> items(b->n);
  uint64_t total = 0;
  for (uint32_t f = 0; f < b->n; ++f) {
    const double want = std::max(1.0, std::floor((double)b->samples[f] * b->sample_scale));
    const uint64_t per_item = N * (uint64_t)c.iters;
    const uint64_t k = std::max<uint64_t>(1, (uint64_t)std::ceil(want / (double)per_item));
    rfk::BatchFrame& F = ft[f];
    memset(&F, 0, sizeof(F));
    F.bins = (float*)b->hist.p + 4 * b->off[f];
    F.W = b->W[f];
    F.H = b->H[f];
    F.item_begin = (uint32_t)total;
    F.n_items = (uint32_t)k;
    rfh::screen_affine(b->desc[f], b->W[f], b->H[f], F.ss);
    items[f] = (uint32_t)k;
    total += k;
  }
  if (total * N >= (1ull << 32)) return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: more than 2^32 particles per launch; raise iters");
  b->n_items = (uint32_t)total;
  HIPCHK(ctx, hipMemcpy(b->frames.p, ft.data(), ft.size() * sizeof(rfk::BatchFrame), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(b->items.p, items.data(), items.size() * 4, hipMemcpyHostToDevice));
  if (b->perm_waves != c.block_waves) {
    const std::vector<uint16_t> tab = rfh::perm_rows(ctx->sp.seed ^ 0xBA7C4ull, c.block_waves);
    HIPCHK(ctx, b->perms.alloc(tab.size() * 2));
    HIPCHK(ctx, hipMemcpy(b->perms.p, tab.data(), tab.size() * 2, hipMemcpyHostToDevice));
    b->perm_waves = c.block_waves;
  }

  rfk::BatchArgs a{};
  a.frames = (const rfk::BatchFrame*)b->frames.p;
  a.blocks = (const Block*)b->blocks.p;
  a.palettes = (const float4*)b->palettes.p;
  a.perms = (const uint16_t*)b->perms.p;
  a.drawn = (unsigned long long*)b->drawn.p;
  a.remaining = (uint32_t*)b->remaining.p;
  a.done_ts = (unsigned long long*)b->done_ts.p;
  a.seed = ctx->sp.seed;
  a.launch = b->launch;
  a.n_frames = b->n;
  a.n_items = b->n_items;
  a.fuse = c.fuse;
  a.n_iter = c.iters;

  if (b->exec) { HIPCHK(ctx, hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
  if (b->graph) { HIPCHK(ctx, hipGraphDestroy(b->graph)); b->graph = nullptr; }
  // Size the persistent grid and compile the flame program (if any) outside
  // the capture.
  const int hm = (c.hist_mode == RF_HIST_LDS && !b->lds_ok) ? RF_HIST_QUAD : c.hist_mode;
  const rfj::Program* prog = nullptr;
  if (ctx->programs && !b->prog_block.empty()) {
    std::string err;
    prog = rfj::get_program(b->prog_block[0], {{c.block_waves, rfj::kBatchKernel + hm}}, err);
    if (!prog) return rfh::set_err(ctx, RF_ERR_HIP, err);
    b->program_compile_s = rfj::program_compile_seconds(prog);
  } else {
    HIPCHK(ctx, rfk::batch_prepare(c.block_waves, hm));
  }
  b->program_used = prog ? 1 : 0;
  HIPCHK(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = rfk::launch_batch_init(a, (float4*)b->hist.p, b->total_bins, (const uint32_t*)b->items.p,
                                        (uint32_t*)b->tiles_done.p, (unsigned long long*)b->t0.p, b->stream);
  if (e == hipSuccess)
    e = prog ? rfj::launch_batch(prog, a, c.block_waves, hm, 0, b->stream)
             : rfk::launch_batch(a, c.block_waves, hm, 0, b->stream);
  for (uint32_t f = 0; f < b->n && e == hipSuccess; ++f) {
    const rf_flame_desc& d = b->desc[f];
    rfk::DEArgs g{};
    g.hist = (const float4*)b->hist.p + b->off[f];
    g.de_out = nullptr;
    g.img_out = (float4*)b->images.p + b->off[f];
    g.W = b->W[f];
    g.H = b->H[f];
    g.R = std::max(0, std::min(d.estimator_radius, 100));  // src/main.cpp:502
    g.rmin = std::max(0, std::min(d.estimator_min, 100));
    g.rad = (signed char*)b->de_rad.p;
    g.wt = (float*)b->de_wt.p;
    g.curve = d.estimator_curve;
    g.do_tonemap = 1;
    g.gamma = d.gamma;
    g.brightness = d.brightness;
    g.vibrancy = d.vibrancy;
    g.k = 1.0f / 10000.0f;                                   // src/main.cpp:228
    g.tiles_done = (uint32_t*)b->tiles_done.p + f;
    g.done_ts = (unsigned long long*)b->de_done_ts.p + f;
    e = rfk::launch_de_tonemap(g, b->stream);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
  if (e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rfh::set_err(ctx, RF_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(e));
  }
  HIPCHK(ctx, e2);
  b->graph = graph;
  HIPCHK(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->ran = false;
  return RF_OK;
}

int run(rf_batch* b, double* ms) {
  rf_ctx* ctx = b->ctx;
  if (!b->exec) return rfh::set_err(ctx, RF_ERR_STATE, "batch: no graph");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(ctx, hipGraphLaunch(b->exec, b->stream));
  HIPCHK(ctx, hipStreamSynchronize(b->stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  b->ran = true;
  return RF_OK;
}

}  // namespace

extern "C" {

int rf_batch_create(rf_ctx* ctx, const rf_frame_desc* frames, uint32_t n, rf_batch** out) {
  if (!ctx || !out) return RF_ERR_ARG;
  *out = nullptr;
  if (!frames || n == 0) return rfh::set_err(ctx, RF_ERR_ARG, "batch: no frames");
  if (n > kMaxFrames) return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: more than 4096 frames");
  rf_batch* b = new rf_batch();
  b->ctx = ctx;
  b->n = n;
  std::vector<Block> blk(n);
  std::vector<float> pal((size_t)n * 256 * 4);
  for (uint32_t f = 0; f < n; ++f) {
    const rf_frame_desc& fr = frames[f];
    if (!fr.flame || fr.width == 0 || fr.height == 0 || fr.samples == 0) {
      delete b;
      return rfh::set_err(ctx, RF_ERR_ARG, "batch: frame " + std::to_string(f) + " has no flame, size or samples");
    }
    std::string err;
    if (!rfh::build_block(*fr.flame, blk[f], err)) {
      delete b;
      return rfh::set_err(ctx, RF_ERR_ARG, "batch: frame " + std::to_string(f) + ": " + err);
    }
    b->desc.push_back(*fr.flame);
    if (!rfh::fixed_point_ok(*fr.flame)) b->lds_ok = false;
    b->W.push_back(fr.width);
    b->H.push_back(fr.height);
    b->samples.push_back(fr.samples);
    b->off.push_back(b->total_bins);
    b->total_bins += (size_t)fr.width * fr.height;
    memcpy(&pal[(size_t)f * 1024], fr.flame->palette, 1024 * sizeof(float));
  }
  int dev = 0, khz = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = b->blocks.alloc(sizeof(Block) * n);
  if (e == hipSuccess) e = b->palettes.alloc(pal.size() * sizeof(float));
  if (e == hipSuccess) e = b->frames.alloc(sizeof(rfk::BatchFrame) * n);
  if (e == hipSuccess) e = b->hist.alloc(b->total_bins * 16);
  if (e == hipSuccess) e = b->images.alloc(b->total_bins * 16);
  if (e == hipSuccess) e = b->drawn.alloc(8ull * n);
  if (e == hipSuccess) e = b->remaining.alloc(4ull * n);
  if (e == hipSuccess) e = b->done_ts.alloc(8ull * n);
  if (e == hipSuccess) e = b->items.alloc(4ull * n);
  if (e == hipSuccess) e = b->tiles_done.alloc(4ull * n);
  if (e == hipSuccess) e = b->de_done_ts.alloc(8ull * n);
  if (e == hipSuccess) e = b->t0.alloc(8);
  size_t max_bins = 0;
  int span = 0;
  for (uint32_t f = 0; f < n; ++f) {
    max_bins = std::max(max_bins, (size_t)b->W[f] * b->H[f]);
    span = std::max(span, std::max(0, std::min(std::max(b->desc[f].estimator_radius, b->desc[f].estimator_min), 100)));
  }
  // frames' DE launches run one after another on the batch's stream, so they
  // share one radius buffer
  if (e == hipSuccess) e = b->de_rad.alloc(max_bins);
  if (e == hipSuccess) e = b->de_wt.alloc(rfk::de_weight_bytes(span));
  if (e == hipSuccess) e = hipMemcpy(b->blocks.p, blk.data(), sizeof(Block) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->palettes.p, pal.data(), pal.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete b;
    return rfh::set_err(ctx, RF_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
  }
  b->clock_khz = (double)khz;
  b->launch = 0x40000000u + ctx->launch++;
  uint32_t big = 0;
  for (uint32_t f = 1; f < n; ++f) if (blk[f].n_xforms > blk[big].n_xforms) big = f;
  bool covers = true;
  for (uint32_t f = 0; f < n && covers; ++f) covers = rfj::structure_covers(blk[big], blk[f]);
  if (covers) b->prog_block.assign(1, blk[big]);
  const int rc = capture(b);
  if (rc != RF_OK) { delete b; return rc; }
  *out = b;
  return RF_OK;
}

void rf_batch_destroy(rf_batch* b) { delete b; }

int rf_batch_set_config(rf_batch* b, const rf_batch_config* cfg) {
  if (!b || !cfg) return RF_ERR_ARG;
  rf_batch_config c = *cfg;
  if (c.fuse == 0) c.fuse = 20;
  const int rc = check_cfg(b->ctx, c);
  if (rc) return rc;
  b->cfg = c;
  return capture(b);
}

int rf_batch_get_config(const rf_batch* b, rf_batch_config* cfg) {
  if (!b || !cfg) return RF_ERR_ARG;
  *cfg = b->cfg;
  return RF_OK;
}

int rf_batch_tune(rf_batch* b, double fraction, rf_batch_config* best_out, double* best_ms_out) {
  if (!b) return RF_ERR_ARG;
  if (!(fraction > 0.0 && fraction <= 1.0)) return rfh::set_err(b->ctx, RF_ERR_ARG, "batch: sample_fraction must be in (0, 1]");
  const rf_batch_config keep = b->cfg;
  const int waves[2] = {4, 8};
  const int modes[2] = {RF_HIST_QUAD, RF_HIST_LDS};
  const uint32_t iters[3] = {64, 256, 1024};
  rf_batch_config best = keep;
  double best_ms = 1e300;
  b->sample_scale = fraction;
  int rc = RF_OK;
  for (int w : waves) {
    for (int m : modes) {
      for (uint32_t it : iters) {
        b->cfg = rf_batch_config{w, m, it, keep.fuse};
        rc = capture(b);
        if (rc == RF_ERR_LIMIT) { rc = RF_OK; continue; }  // too many particles at this iters
        double ms = 0.0;
        if (rc == RF_OK) rc = run(b, nullptr);   // warm
        if (rc == RF_OK) rc = run(b, &ms);
        if (rc != RF_OK) break;
        if (ms < best_ms) { best_ms = ms; best = b->cfg; }
      }
      if (rc != RF_OK) break;
    }
    if (rc != RF_OK) break;
  }
  b->sample_scale = 1.0;
  b->cfg = rc == RF_OK ? best : keep;
  const int rc2 = capture(b);
  if (rc != RF_OK) return rc;
  if (rc2 != RF_OK) return rc2;
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return RF_OK;
}

int rf_batch_program_info(const rf_batch* b, double* compile_seconds) {
  if (!b) return RF_ERR_ARG;
  if (compile_seconds) *compile_seconds = b->program_compile_s;
  return b->program_used;
}

int rf_batch_run(rf_batch* b, double* ms) {
  if (!b) return RF_ERR_ARG;
  return run(b, ms);
}

int rf_batch_frame_stats(const rf_batch* b, uint64_t* drawn, double* chaos_ms, double* image_ms) {
  if (!b) return RF_ERR_ARG;
  rf_ctx* ctx = b->ctx;
  if (!b->ran) return rfh::set_err(ctx, RF_ERR_STATE, "batch: not run since the last configuration");
  std::vector<unsigned long long> d(b->n), cd(b->n), id(b->n);
  unsigned long long t0 = 0;
  HIPCHK(ctx, hipMemcpy(d.data(), b->drawn.p, 8ull * b->n, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(cd.data(), b->done_ts.p, 8ull * b->n, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(id.data(), b->de_done_ts.p, 8ull * b->n, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(&t0, b->t0.p, 8, hipMemcpyDeviceToHost));
  const double k = b->clock_khz > 0 ? 1.0 / b->clock_khz : 0.0;  // ms per tick
  for (uint32_t f = 0; f < b->n; ++f) {
    if (drawn) drawn[f] = d[f];
    if (chaos_ms) chaos_ms[f] = (double)(long long)(cd[f] - t0) * k;
    if (image_ms) image_ms[f] = (double)(long long)(id[f] - t0) * k;
  }
  return RF_OK;
}

void* rf_batch_hist_ptr(rf_batch* b, uint32_t i) {
  return (b && i < b->n) ? (void*)((float4*)b->hist.p + b->off[i]) : nullptr;
}

void* rf_batch_image_ptr(rf_batch* b, uint32_t i) {
  return (b && i < b->n) ? (void*)((float4*)b->images.p + b->off[i]) : nullptr;
}

int rf_batch_download(rf_batch* b, uint32_t i, int32_t which, float* host) {
  if (!b || !host) return RF_ERR_ARG;
  if (i >= b->n || (which != 0 && which != 1)) return rfh::set_err(b->ctx, RF_ERR_ARG, "batch: bad frame or buffer");
  const void* src = which == 0 ? rf_batch_hist_ptr(b, i) : rf_batch_image_ptr(b, i);
  HIPCHK(b->ctx, hipMemcpy(host, src, (size_t)b->W[i] * b->H[i] * 16, hipMemcpyDeviceToHost));
  return RF_OK;
}

}  // extern "C"
