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
  // A replay runs in slices: ranges of work items in frame order, each one
  // chaos-game launch (plus, with RF_HIST_SORT, its record passes), followed
  // by the density estimation of the frames it completes on a second stream.
  struct Slice {
    uint32_t item0 = 0, n_items = 0;
    uint32_t span_base = 0, span_bins = 0;  // RF_HIST_SORT: bins the slice's records address
    uint32_t n_buckets = 0, n_tiles = 0, n_chunks = 0, n_segs = 0;
    size_t meta = 0;                         // first word of the slice's capacities and work lists
  };
  std::vector<Slice> slices;
  std::vector<hipEvent_t> events;
  hipStream_t de_stream = nullptr;
  // RF_HIST_SORT: record buffers and cursors shared by the slices (they run
  // one after another), per slice capacities, work lists and tile palettes
  DevBuf cbuf, tbuf, bcur, tcur, sort_meta, tile_mass;
  std::vector<uint32_t> tile_frame;  // per 2^12-bin tile of the batch histogram: its frame, ~0u for padding
  bool sort_ok = true;               // every frame has unit opacities and a palette in [0, 1]
  bool have_hist = false;            // the histograms hold a finished replay (capacity estimates)
  int used_mode = RF_HIST_LDS;       // histogram mode the captured graph runs
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
    if (de_stream) { (void)hipStreamSynchronize(de_stream); (void)hipStreamDestroy(de_stream); }
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
};

namespace {

constexpr uint32_t kMaxFrames = 4096;

int check_cfg(rf_ctx* ctx, const rf_batch_config& c) {
  if (c.block_waves != 4 && c.block_waves != 8) return rfh::set_err(ctx, RF_ERR_ARG, "batch: block_waves must be 4 or 8");
  if (c.hist_mode != RF_HIST_QUAD && c.hist_mode != RF_HIST_LDS && c.hist_mode != RF_HIST_SORT)
    return rfh::set_err(ctx, RF_ERR_ARG, "batch: hist_mode must be RF_HIST_QUAD, RF_HIST_LDS or RF_HIST_SORT");
  if (c.iters == 0 || c.iters > (1u << 20)) return rfh::set_err(ctx, RF_ERR_ARG, "batch: iters must be in 1..2^20");
  if (c.fuse > 1024) return rfh::set_err(ctx, RF_ERR_ARG, "batch: fuse must be at most 1024");
  return RF_OK;
}

constexpr uint64_t kSliceSamples = 1ull << 29;   // samples per slice at most
constexpr uint32_t kTileBins = 1u << rfk::SORT_TILE_SHIFT;
constexpr uint32_t kBucketBins = 1u << rfk::SORT_COARSE_SHIFT;
constexpr uint32_t kMaxSpan = (uint32_t)rfk::SORT_MAX_BUCKETS << rfk::SORT_COARSE_SHIFT;

// Splits the items (numbered frame by frame) into slices of at most
// kSliceSamples samples whose frames span at most 2^27 bins (counted from
// the bucket boundary at or below the first frame's first bin).
void plan_slices(rf_batch* b, const std::vector<rfk::BatchFrame>& ft, uint64_t per_item) {
  b->slices.clear();
  const uint32_t max_items = (uint32_t)std::max<uint64_t>(1, kSliceSamples / per_item);
  uint32_t f = 0, item = 0;
  while (item < b->n_items) {
    while (ft[f].item_begin + ft[f].n_items <= item) ++f;
    rf_batch::Slice sl;
    sl.item0 = item;
    sl.span_base = (uint32_t)(b->off[f] / kBucketBins * kBucketBins);
    uint32_t end = item, g = f;
    while (end < b->n_items && end - item < max_items) {
      while (ft[g].item_begin + ft[g].n_items <= end) ++g;
      const uint64_t span_end = (b->off[g] + (uint64_t)b->W[g] * b->H[g] + kTileBins - 1) / kTileBins * kTileBins;
      if (g > f && span_end - sl.span_base > kMaxSpan) break;
      end = std::min<uint32_t>(ft[g].item_begin + ft[g].n_items, item + max_items);
    }
    const uint32_t last = end - 1;
    uint32_t fl = f;
    while (ft[fl].item_begin + ft[fl].n_items <= last) ++fl;
    const uint64_t span_end = (b->off[fl] + (uint64_t)b->W[fl] * b->H[fl] + kTileBins - 1) / kTileBins * kTileBins;
    sl.span_bins = (uint32_t)(span_end - sl.span_base);
    sl.n_items = end - item;
    b->slices.push_back(sl);
    item = end;
  }
}

// RF_HIST_SORT: capacities and work lists of every slice.  A frame's samples
// in a slice (its items there x samples per item) are spread over its tiles
// in proportion to the count channel of the last finished replay (uniformly,
// at three times the mean, when there is none); 20 % headroom plus a fixed
// reserve per tile and bucket, and records that still do not fit go to the
// histogram with quad atomics, so a poor estimate costs speed only.
int plan_sort(rf_batch* b, const std::vector<rfk::BatchFrame>& ft, uint64_t per_item) {
  rf_ctx* ctx = b->ctx;
  const uint32_t n_tiles_all = (uint32_t)(b->total_bins / kTileBins);
  std::vector<double> w(n_tiles_all, 0.0);  // fraction of its frame's samples per tile
  std::vector<float> mass;
  if (b->have_hist) {
    mass.resize(n_tiles_all);
    if (b->tile_mass.bytes < (size_t)n_tiles_all * 4) HIPCHK(ctx, b->tile_mass.alloc((size_t)n_tiles_all * 4));
    HIPCHK(ctx, rfk::launch_sort_tile_mass((const float4*)b->hist.p, (uint32_t)b->total_bins, n_tiles_all,
                                           (float*)b->tile_mass.p, b->stream));
    HIPCHK(ctx, hipMemcpyAsync(mass.data(), b->tile_mass.p, (size_t)n_tiles_all * 4, hipMemcpyDeviceToHost,
                               b->stream));
    HIPCHK(ctx, hipStreamSynchronize(b->stream));
  }
  for (uint32_t f = 0; f < b->n; ++f) {
    const uint32_t t0 = (uint32_t)(b->off[f] / kTileBins);
    const uint32_t nt = (uint32_t)(((uint64_t)b->W[f] * b->H[f] + kTileBins - 1) / kTileBins);
    double tot = 0.0;
    if (!mass.empty()) for (uint32_t t = 0; t < nt; ++t) tot += mass[t0 + t];
    for (uint32_t t = 0; t < nt; ++t) w[t0 + t] = tot > 0.0 ? mass[t0 + t] / tot : 3.0 / nt;
  }
  std::vector<uint32_t> meta;
  uint64_t max_cb = 0, max_ct = 0;
  for (auto& sl : b->slices) {
    const uint32_t nb = (sl.span_bins + kBucketBins - 1) / kBucketBins;
    const uint32_t nt = nb * (uint32_t)rfk::SORT_TILES_PER_BUCKET;
    const uint32_t gt0 = sl.span_base / kTileBins;
    std::vector<double> e(nt, 0.0);     // expected records per tile
    std::vector<char> active(nt, 0);    // the tile belongs to a frame with samples in the slice
    for (uint32_t f = 0; f < b->n; ++f) {
      const uint32_t i0 = std::max(ft[f].item_begin, sl.item0);
      const uint32_t i1 = std::min(ft[f].item_begin + ft[f].n_items, sl.item0 + sl.n_items);
      if (i0 >= i1) continue;
      const double sf = (double)(i1 - i0) * (double)per_item;
      const uint32_t t0 = (uint32_t)(b->off[f] / kTileBins);
      const uint32_t ntf = (uint32_t)(((uint64_t)b->W[f] * b->H[f] + kTileBins - 1) / kTileBins);
      for (uint32_t t = 0; t < ntf; ++t) {
        const uint32_t lt = t0 + t - gt0;
        e[lt] += sf * w[t0 + t];
        active[lt] = 1;
      }
    }
    sl.meta = meta.size();
    sl.n_buckets = nb;
    sl.n_tiles = nt;
    const size_t base = meta.size();
    meta.resize(base + 2 * (size_t)nb + 3 * (size_t)nt, 0u);  // boff bcap toff tcap tile_pal
    uint32_t* boff = meta.data() + base;
    uint32_t* bcap = boff + nb;
    uint32_t* toff = bcap + nb;
    uint32_t* tcap = toff + nt;
    uint32_t* tpal = tcap + nt;
    uint64_t cb = 0, ct = 0;
    for (uint32_t t = 0; t < nt; ++t) {
      uint64_t c = active[t] ? (uint64_t)std::ceil(e[t] * 1.2) + 256 : 0;
      c = (c + 3) & ~3ull;  // every tile's records start 16-byte aligned
      toff[t] = (uint32_t)ct;
      tcap[t] = (uint32_t)std::min<uint64_t>(c, 0x7ffffffcull);
      ct += tcap[t];
      // the slice's last bucket may reach past the batch histogram's end
      const uint32_t fr = gt0 + t < b->tile_frame.size() ? b->tile_frame[gt0 + t] : ~0u;
      tpal[t] = fr == ~0u ? 0u : fr * 256u;
    }
    for (uint32_t k = 0; k < nb; ++k) {
      double eb = 0.0;
      bool any = false;
      for (uint32_t t = k * rfk::SORT_TILES_PER_BUCKET; t < (k + 1) * rfk::SORT_TILES_PER_BUCKET; ++t) {
        eb += e[t];
        any = any || active[t];
      }
      uint64_t c = any ? (uint64_t)std::ceil(eb * 1.2) + 2048 : 0;
      c = (c + 3) & ~3ull;
      boff[k] = (uint32_t)cb;
      bcap[k] = (uint32_t)std::min<uint64_t>(c, 0x7ffffffcull);
      cb += bcap[k];
    }
    if (cb >= 0xffffffffull || ct >= 0xffffffffull)
      return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: more than 2^32 records in a slice");
    max_cb = std::max(max_cb, cb);
    max_ct = std::max(max_ct, ct);
    std::vector<uint32_t> chunks;
    for (uint32_t k = 0; k < nb; ++k)
      for (uint32_t q = 0; (uint64_t)q * rfk::SORT_CHUNK < bcap[k]; ++q) {
        if (q > 0xffffu) return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: sort bucket too large");
        chunks.push_back((k << 16) | q);
      }
    std::vector<std::pair<uint32_t, uint32_t>> segs;  // (records the segment can hold, job)
    for (uint32_t t = 0; t < nt; ++t) {
      if (tcap[t] == 0) continue;
      if (tcap[t] <= (uint32_t)rfk::SORT_SEG16) { segs.emplace_back(tcap[t], t << 16); continue; }
      for (uint32_t q = 0; (uint64_t)q * rfk::SORT_SEG15 < tcap[t]; ++q) {
        if (q > 0xffffu) return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: sort tile too large");
        segs.emplace_back(std::min<uint32_t>(tcap[t] - q * (uint32_t)rfk::SORT_SEG15, rfk::SORT_SEG15), (t << 16) | q);
      }
    }
    rfh::order_segments(segs);
    sl.n_chunks = (uint32_t)chunks.size();
    sl.n_segs = (uint32_t)segs.size();
    meta.insert(meta.end(), chunks.begin(), chunks.end());
    for (const auto& sg : segs) meta.push_back(sg.second);
  }
  if (b->cbuf.bytes < max_cb * 4) HIPCHK(ctx, b->cbuf.alloc(max_cb * 4));
  if (b->tbuf.bytes < max_ct * 4) HIPCHK(ctx, b->tbuf.alloc(max_ct * 4));
  if (!b->bcur.p) HIPCHK(ctx, b->bcur.alloc((size_t)rfk::SORT_MAX_BUCKETS * 4));
  if (!b->tcur.p) HIPCHK(ctx, b->tcur.alloc((size_t)rfk::SORT_MAX_BUCKETS * rfk::SORT_TILES_PER_BUCKET * 4));
  if (b->sort_meta.bytes < meta.size() * 4) HIPCHK(ctx, b->sort_meta.alloc(meta.size() * 4));
  HIPCHK(ctx, hipMemcpy(b->sort_meta.p, meta.data(), meta.size() * 4, hipMemcpyHostToDevice));
  return RF_OK;
}

rfk::SortArgs slice_sort_args(rf_batch* b, const rf_batch::Slice& sl) {
  rfk::SortArgs a{};
  const uint32_t* m = (const uint32_t*)b->sort_meta.p + sl.meta;
  const uint32_t nb = sl.n_buckets, nt = sl.n_tiles;
  a.cbuf = (uint32_t*)b->cbuf.p;
  a.tbuf = (uint32_t*)b->tbuf.p;
  a.bcur = (uint32_t*)b->bcur.p;
  a.tcur = (uint32_t*)b->tcur.p;
  a.boff = m;
  a.bcap = m + nb;
  a.toff = m + 2 * (size_t)nb;
  a.tcap = m + 2 * (size_t)nb + nt;
  a.tile_pal = m + 2 * (size_t)nb + 2 * (size_t)nt;
  a.chunks = m + 2 * (size_t)nb + 3 * (size_t)nt;
  a.segs = a.chunks + sl.n_chunks;
  a.n_buckets = nb;
  a.n_tiles = nt;
  a.n_chunks = sl.n_chunks;
  a.n_segs = sl.n_segs;
  return a;
}

// Lays out the work items for the current config and sample scale, plans the
// slices, uploads the frame table and captures the graph.
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
    F.goff = (uint32_t)b->off[f];
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
  int hm = c.hist_mode;
  if (hm == RF_HIST_SORT && !b->sort_ok) hm = RF_HIST_LDS;
  if (hm == RF_HIST_LDS && !b->lds_ok) hm = RF_HIST_QUAD;
  const uint64_t per_item = N * (uint64_t)c.iters;
  plan_slices(b, ft, per_item);
  if (hm == RF_HIST_SORT) {
    const int rc = plan_sort(b, ft, per_item);
    if (rc) return rc;
  }
  while (b->events.size() < b->slices.size() + 1) {
    hipEvent_t e = nullptr;
    HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->events.push_back(e);
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
  a.item0 = 0;
  a.n_items = b->n_items;
  a.fuse = c.fuse;
  a.n_iter = c.iters;

  if (b->exec) { HIPCHK(ctx, hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
  if (b->graph) { HIPCHK(ctx, hipGraphDestroy(b->graph)); b->graph = nullptr; }
  // Size the persistent grid and compile the flame program (if any) outside
  // the capture.
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
  b->used_mode = hm;
  HIPCHK(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = rfk::launch_batch_init(a, (float4*)b->hist.p, b->total_bins, (const uint32_t*)b->items.p,
                                        (uint32_t*)b->tiles_done.p, (unsigned long long*)b->t0.p, b->stream);
  uint32_t next_frame = 0;   // the first frame whose density estimation is not yet captured
  size_t n_forks = 0;
  for (size_t k = 0; k < b->slices.size() && e == hipSuccess; ++k) {
    const rf_batch::Slice& sl = b->slices[k];
    rfk::BatchArgs as = a;
    as.item0 = sl.item0;
    as.n_items = sl.n_items;
    float* span = (float*)b->hist.p + 4 * (size_t)sl.span_base;
    if (hm == RF_HIST_SORT) {
      as.span_base = sl.span_base;
      as.span_bins = span;
      as.sort = slice_sort_args(b, sl);
      e = rfk::launch_sort_reset(as.sort, b->stream);
    }
    if (e == hipSuccess)
      e = prog ? rfj::launch_batch(prog, as, c.block_waves, hm, 0, b->stream)
               : rfk::launch_batch(as, c.block_waves, hm, 0, b->stream);
    if (e == hipSuccess && hm == RF_HIST_SORT) e = rfk::launch_sort_rebucket(as.sort, span, a.palettes, b->stream);
    if (e == hipSuccess && hm == RF_HIST_SORT)
      e = rfk::launch_sort_tiles(as.sort, span, sl.span_bins, a.palettes, b->stream);
    // frames whose last item ran in this slice: their density estimation
    // and tonemap on the second stream, overlapping the next slice
    const uint32_t end = sl.item0 + sl.n_items;
    uint32_t done_to = next_frame;
    while (done_to < b->n && ft[done_to].item_begin + ft[done_to].n_items <= end) ++done_to;
    if (e == hipSuccess && done_to > next_frame) {
      e = hipEventRecord(b->events[k], b->stream);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->de_stream, b->events[k], 0);
      ++n_forks;
    }
    for (uint32_t f = next_frame; f < done_to && e == hipSuccess; ++f) {
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
      e = rfk::launch_de_tonemap(g, b->de_stream);
    }
    next_frame = done_to;
  }
  // join the second stream before the capture ends
  if (e == hipSuccess && n_forks) {
    e = hipEventRecord(b->events[b->slices.size()], b->de_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->events[b->slices.size()], 0);
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
  b->have_hist = true;
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
    // the sort path's records carry no weight
    for (int x = 0; x < fr.flame->n_xforms; ++x) if (fr.flame->xforms[x].opacity != 1.0f) b->sort_ok = false;
    if (fr.flame->has_final && fr.flame->final_xform.opacity != 1.0f) b->sort_ok = false;
    b->W.push_back(fr.width);
    b->H.push_back(fr.height);
    b->samples.push_back(fr.samples);
    // frames start on 2^12-bin tile boundaries: every tile of the sort path
    // belongs to one frame (and palette)
    b->off.push_back(b->total_bins);
    const size_t nbins = (size_t)fr.width * fr.height;
    const size_t ntiles = (nbins + kTileBins - 1) / kTileBins;
    for (size_t t = 0; t < ntiles; ++t) b->tile_frame.push_back(f);
    b->total_bins += ntiles * kTileBins;
    memcpy(&pal[(size_t)f * 1024], fr.flame->palette, 1024 * sizeof(float));
  }
  if (!b->lds_ok) b->sort_ok = false;
  if (b->total_bins >= (1ull << 32)) {
    delete b;
    return rfh::set_err(ctx, RF_ERR_LIMIT, "batch: more than 2^32 bins in all frames");
  }
  int dev = 0, khz = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->de_stream, hipStreamNonBlocking);
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
  // quad and LDS first: their replays give the sort path its capacity estimates
  const int modes[3] = {RF_HIST_QUAD, RF_HIST_LDS, RF_HIST_SORT};
  const uint32_t iters[3] = {64, 256, 1024};
  rf_batch_config best = keep;
  double best_ms = 1e300;
  b->sample_scale = fraction;
  int rc = RF_OK;
  for (int w : waves) {
    for (int m : modes) {
      if (m == RF_HIST_SORT && !b->sort_ok) continue;
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
