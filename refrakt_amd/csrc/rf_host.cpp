// Host library: flame model, simulation state and the C ABI entry points.
//
// The orchestration mirrors src/flame.cpp (set_sim_parameters, warmup,
// draw_to_bins) and the inline DE/tonemap passes of src/main.cpp:490-535;
// the device work is done by the kernels of rf_k_compat.hip, rf_k_fast.hip and
// rf_k_misc.hip on one HIP stream per context.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "rf_host.hpp"
#include "rf_jit.hpp"
#include "rf_internal.hpp"

#include "gen/variation_table.inc"

using rfd::Block;

namespace rfh {

hipError_t DevBuf::alloc(size_t n) {
  if (p && bytes == n) return hipSuccess;
  release();
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) { p = nullptr; bytes = 0; return e; }
  bytes = n;
  return hipSuccess;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

hipError_t HostBuf::alloc(size_t n) {
  if (p && bytes >= n) return hipSuccess;
  release();
  if (n == 0) return hipSuccess;
  hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
  if (e != hipSuccess) { p = nullptr; bytes = 0; return e; }
  e = hipHostGetDevicePointer(&dev, p, 0);
  if (e != hipSuccess) { release(); return e; }
  bytes = n;
  return hipSuccess;
}

void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  dev = nullptr;
  bytes = 0;
}

void order_segments(std::vector<std::pair<uint32_t, uint32_t>>& segs) {
  constexpr int kStep = 10;
  constexpr uint32_t kClasses = ((uint32_t)rfk::SORT_SEG15 >> kStep) + 1;
  static_assert(rfk::SORT_SEG16 <= rfk::SORT_SEG15, "segment sizes");
  uint32_t start[kClasses + 1] = {0};
  // class 0 holds the largest segments
  auto cls = [](uint32_t n) { return kClasses - 1 - std::min<uint32_t>(n >> kStep, kClasses - 1); };
  for (const auto& sg : segs) ++start[cls(sg.first) + 1];
  for (uint32_t k = 0; k < kClasses; ++k) start[k + 1] += start[k];
  std::vector<std::pair<uint32_t, uint32_t>> out(segs.size());
  for (const auto& sg : segs) out[start[cls(sg.first)]++] = sg;
  segs.swap(out);
}

const rf_var_info* var_info(int id) {
  if (id < 0 || id >= RF_GEN_NUM_VARIATIONS) return nullptr;
  return &k_var_table[id];
}

namespace {
struct ParamIndex {
  std::unordered_map<std::string, std::pair<int, int>> params;
  std::unordered_map<std::string, int> vars;
  ParamIndex() {
    for (int v = 0; v < RF_GEN_NUM_VARIATIONS; ++v) {
      vars[k_var_table[v].name] = v;
      for (int k = 0; k < k_var_table[v].n_params; ++k) params[k_var_table[v].params[k]] = {v, k};
    }
  }
};
const ParamIndex& pindex() {
  static ParamIndex idx;
  return idx;
}

// splitmix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA'14); drives the shuffle tables and row schedule.
uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint32_t bounded(uint64_t& s, uint32_t n) {
  return (uint32_t)(((uint64_t)(uint32_t)(splitmix64(s) >> 32) * n) >> 32);
}
}  // namespace

bool param_lookup(const char* name, int* var, int* k) {
  auto& idx = pindex();
  auto it = idx.params.find(name);
  if (it == idx.params.end()) return false;
  *var = it->second.first;
  *k = it->second.second;
  return true;
}

// Random permutation of [0, n) (Fisher-Yates), this library's stand-in for the
// reference's std::shuffle rows (src/shuffle_buffers.cpp:9-21).
void shuffle_row(uint32_t* row, uint32_t n, uint64_t seed) {
  uint64_t s = seed;
  for (uint32_t i = 0; i < n; ++i) row[i] = i;
  for (uint32_t i = n - 1; i > 0; --i) {
    uint32_t j = bounded(s, i + 1);
    std::swap(row[i], row[j]);
  }
}

static bool name_less(int a, int b) { return strcmp(k_var_table[a].name, k_var_table[b].name) < 0; }

bool build_block(const rf_flame_desc& d, Block& b, std::string& err) {
  memset(&b, 0, sizeof(b));
  if (d.n_xforms < 1 || d.n_xforms > rfd::MAX_XF) { err = "n_xforms out of range"; return false; }
  b.n_xforms = (uint32_t)d.n_xforms;
  b.has_final = d.has_final ? 1u : 0u;
  // copy_flame_data_to_buffer (src/flame.cpp:78-82): weights normalised by
  // the float sum over the non-final xforms; get_xform_id sums them in order.
  float total = 0.0f;
  for (int i = 0; i < d.n_xforms; ++i) total += d.xforms[i].weight;
  if (!(total > 0.0f)) { err = "xform weights must sum to a positive value"; return false; }
  float sum = 0.0f;
  for (int i = 0; i < d.n_xforms; ++i) {
    const float wn = d.xforms[i].weight / total;
    sum = (i == 0) ? wn : sum + wn;
    b.cdf[i] = sum;
    // Largest raw draw x with float(x) * 2^-32 <= cdf[i] (the float
    // conversion rounds to nearest even on host and device alike, and is
    // monotone in x, so a binary search finds it).
    auto u_of = [](uint32_t x) { return (float)x * 0x1p-32f; };
    if (!(sum >= 0.0f)) {
      b.never_mask |= 1u << i;
    } else if (u_of(0xffffffffu) <= sum) {
      b.cdf_lim[i] = 0xffffffffu;
    } else {
      uint32_t lo = 0, hi = 0xffffffffu;  // u_of(lo) <= sum < u_of(hi)
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (u_of(mid) <= sum) lo = mid; else hi = mid;
      }
      b.cdf_lim[i] = lo;
    }
  }
  uint32_t nt = 0;
  auto put = [&](const rf_xform& x, rfd::Xform& X) -> bool {
    for (int k = 0; k < 6; ++k) { X.aff[k] = x.affine[k]; X.post[k] = x.post[k]; }
    X.flags = x.has_post ? rfd::XF_HAS_POST : 0u;
    X.color = x.color;
    X.color_speed = x.color_speed;
    X.opacity = x.opacity;
    // rf_mix's factors in single precision (no contraction: the same
    // roundings as on the device)
    const float keep = 1.0f - x.color_speed, add = x.color * x.color_speed;
    X.c_keep = keep;
    X.c_add = add;
    if (x.n_vars < 0 || x.n_vars > RF_MAX_VARS_PER_XFORM) { err = "n_vars out of range"; return false; }
    // std::map<std::string,float> semantics: one entry per name (last wins),
    // iterated in name order (src/flame.hpp:26).
    std::vector<int> order;
    for (int t = 0; t < x.n_vars; ++t) {
      const int id = x.vars[t].id;
      if (!var_info(id)) { err = "unknown variation id " + std::to_string(id); return false; }
      bool dup = false;
      for (int q = t + 1; q < x.n_vars; ++q) if (x.vars[q].id == id) dup = true;
      if (!dup) order.push_back(t);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return name_less(x.vars[a].id, x.vars[c].id); });
    std::vector<int> pre, reg;
    for (int t : order) ((k_var_table[x.vars[t].id].flags & rfd::TF_PRE_XFORM) ? pre : reg).push_back(t);
    X.term_begin = nt;
    X.n_pre = (uint32_t)pre.size();
    X.n_terms = (uint32_t)(pre.size() + reg.size());
    X.precalc = 0;
    if (nt + X.n_terms > (uint32_t)rfd::MAX_TERMS) { err = "too many variation terms in the flame"; return false; }
    std::vector<int> all(pre);
    all.insert(all.end(), reg.begin(), reg.end());
    for (int t : all) {
      const rf_var_term& v = x.vars[t];
      rfd::Term& T = b.terms[nt++];
      T.var = (uint32_t)v.id;
      T.flags = k_var_table[v.id].flags & (rfd::TF_NO_WEIGHT_MUL | rfd::TF_PRE_XFORM);
      T.weight = v.weight;
      for (int k = 0; k < rfd::MAX_PRM; ++k) T.prm[k] = (k < k_var_table[v.id].n_params) ? v.params[k] : 0.0f;
      if (!(T.flags & rfd::TF_PRE_XFORM)) X.precalc |= k_var_table[v.id].precalc;
    }
    return true;
  };
  for (int i = 0; i < d.n_xforms; ++i) if (!put(d.xforms[i], b.xf[i])) return false;
  if (d.has_final && !put(d.final_xform, b.xf[rfd::FINAL])) return false;
  b.n_terms = nt;
  return true;
}

std::vector<float> reference_param_block(const rf_flame_desc& d) {
  std::vector<float> out;
  float normal_weight = 0.0f;
  for (int i = 0; i < d.n_xforms; ++i) normal_weight += d.xforms[i].weight;
  auto push = [&](const rf_xform& x) {
    out.push_back(x.weight / normal_weight);
    for (int a = 0; a < 6; ++a) out.push_back(x.affine[a]);
    if (x.has_post) for (int a = 0; a < 6; ++a) out.push_back(x.post[a]);
    std::vector<std::pair<std::string, float>> vars, prms;
    for (int t = 0; t < x.n_vars; ++t) {
      const rf_var_info* vi = var_info(x.vars[t].id);
      if (!vi) continue;
      vars.emplace_back(vi->name, x.vars[t].weight);
      for (int k = 0; k < vi->n_params; ++k) prms.emplace_back(vi->params[k], x.vars[t].params[k]);
    }
    auto by_name = [](const std::pair<std::string, float>& a, const std::pair<std::string, float>& b) { return a.first < b.first; };
    std::stable_sort(vars.begin(), vars.end(), by_name);
    std::stable_sort(prms.begin(), prms.end(), by_name);
    for (auto& v : vars) out.push_back(v.second);
    for (auto& v : prms) out.push_back(v.second);
    out.push_back(x.color);
    out.push_back(x.color_speed);
    out.push_back(x.opacity);
    out.push_back(x.rotation_frequency);
  };
  for (int i = 0; i < d.n_xforms; ++i) push(d.xforms[i]);
  if (d.has_final) push(d.final_xform);
  return out;
}

void screen_affine(const rf_flame_desc& d, uint32_t w, uint32_t h, float out[6]) {
  float b[6] = {1, 0, 0, 1, 0, 0};
  auto translate = [](float a[6], float t0, float t1) {
    const float r4 = a[0] * t0 + a[2] * t1 + a[4];
    const float r5 = a[1] * t0 + a[3] * t1 + a[5];
    a[4] = r4;
    a[5] = r5;
  };
  translate(b, (float)w / 2.0f, (float)h / 2.0f);
  const float s = d.scale * (float)h / (float)d.size[1];
  for (int k = 0; k < 4; ++k) b[k] = b[k] * s;
  const float rad = 0.01745329251f * d.rotate;
  const float sino = sinf(rad), coso = cosf(rad);
  const float r0 = b[0] * coso + b[2] * sino;
  const float r1 = b[1] * coso + b[3] * sino;
  const float r2 = b[2] * coso - b[0] * sino;
  const float r3 = b[3] * coso - b[1] * sino;
  b[0] = r0; b[1] = r1; b[2] = r2; b[3] = r3;
  translate(b, -d.center[0], -d.center[1]);
  for (int k = 0; k < 6; ++k) out[k] = b[k];
}

}  // namespace rfh

using namespace rfh;

// ---------------------------------------------------------------- helpers
int rfh::set_err(rf_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

bool rfh::fixed_point_ok(const rf_flame_desc& d) {
  for (int i = 0; i < RF_PALETTE_SIZE; ++i)
    for (int c = 0; c < 3; ++c)
      if (!(d.palette[i][c] >= 0.0f && d.palette[i][c] <= 1.0f)) return false;
  for (int i = 0; i < d.n_xforms; ++i)
    if (!(d.xforms[i].opacity >= 0.0f && d.xforms[i].opacity <= 1.0f)) return false;
  if (d.has_final && !(d.final_xform.opacity >= 0.0f && d.final_xform.opacity <= 1.0f)) return false;
  return true;
}

std::vector<uint16_t> rfh::perm_rows(uint64_t seed, int waves) {
  const uint32_t N = 64u * (uint32_t)waves;
  std::vector<uint16_t> tab((size_t)rfk::NPERM * N);
  std::vector<uint32_t> row(N);
  for (int r = 0; r < rfk::NPERM; ++r) {
    shuffle_row(row.data(), N, seed ^ 0x5EED5EED00000000ull ^ (uint64_t)(r + 1) * 0x2545F4914F6CDD1Dull);
    for (uint32_t i = 0; i < N; ++i) tab[(size_t)r * N + i] = (uint16_t)row[i];
  }
  return tab;
}

static int validate_sim(rf_ctx* ctx, uint64_t P, uint32_t T) {
  if (T == 0) return set_err(ctx, RF_ERR_ARG, "temporal_samples must be positive");
  if (P == 0 || P % (256ull * T) != 0)
    return set_err(ctx, RF_ERR_ARG, "total_particles must be a positive multiple of 256 * temporal_samples");
  // particle and batch indices are 32-bit in both engines
  if (P > 0xffffffffull) return set_err(ctx, RF_ERR_LIMIT, "total_particles must be below 2^32");
  return RF_OK;
}

static int ensure_compat(rf_ctx* ctx) {
  if (ctx->compat_ready) return RF_OK;
  const uint64_t P = ctx->sp.total_particles;
  const uint64_t ppt = ctx->ppt;
  HIPCHK(ctx, ctx->swap.alloc(P * 16));
  HIPCHK(ctx, ctx->rs.alloc(P * 16));
  HIPCHK(ctx, rfk::launch_init_rs((uint4*)ctx->rs.p, P, ctx->stream));
  // shuffle rows (set_sim_parameters, src/flame.cpp:114-130)
  const uint32_t R = ctx->shuffle_count;
  std::vector<uint32_t> rows((size_t)R * ppt);
  for (uint32_t r = 0; r < R; ++r)
    shuffle_row(rows.data() + (size_t)r * ppt, (uint32_t)ppt, ctx->sp.seed ^ (0xA5A5A5A5ull * (r + 1)));
  HIPCHK(ctx, ctx->shuf.alloc(rows.size() * 4));
  HIPCHK(ctx, hipMemcpyAsync(ctx->shuf.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  // Hammersley sample points (src/hammersley.cpp:29-48)
  std::vector<float> pts(ppt * 4, 0.0f);
  {
    const uint32_t cnt = (uint32_t)ppt;
    uint32_t mx = cnt;
    if (cnt % 2 != 0) { uint32_t v = cnt - 1; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; mx = v + 1; }
    const float inv_max = 1.0f / (float)mx;
    uint32_t r = 0, v = mx;
    while (v >>= 1) ++r;
    for (uint32_t i = 0; i < cnt; ++i) {
      uint32_t n = i;
      n = (n << 16) | (n >> 16);
      n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
      n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
      n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
      n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
      const uint32_t fl = r ? (n >> (32 - r)) : 0u;
      pts[4 * (size_t)i + 0] = ((float)i * inv_max) * 2.0f - 1.0f;
      pts[4 * (size_t)i + 1] = ((float)fl * inv_max) * 2.0f - 1.0f;
    }
  }
  HIPCHK(ctx, ctx->samples.alloc(pts.size() * 4));
  HIPCHK(ctx, hipMemcpyAsync(ctx->samples.p, pts.data(), pts.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->compat_ready = true;
  return RF_OK;
}

static int fast_waves(const rf_ctx* ctx) {
  // 8 waves suit the LDS bin cache (its 2048-slot table shares the CU's LDS
  // between two workgroups) and cost the direct modes about 1 %.
  const int w = ctx->sp.block_waves;
  return w == 4 ? 4 : 8;
}

static int ensure_perms(rf_ctx* ctx) {
  const int waves = fast_waves(ctx);
  if (ctx->perms.p && ctx->perm_waves == waves) return RF_OK;
  const std::vector<uint16_t> tab = perm_rows(ctx->sp.seed, waves);
  HIPCHK(ctx, ctx->perms.alloc(tab.size() * 2));
  HIPCHK(ctx, hipMemcpy(ctx->perms.p, tab.data(), tab.size() * 2, hipMemcpyHostToDevice));
  ctx->perm_waves = waves;
  return RF_OK;
}

static void next_rows(rf_ctx* ctx, uint32_t n_pairs, std::vector<uint32_t>& rows) {
  rows.resize(2 * (size_t)n_pairs);
  for (auto& r : rows) r = bounded(ctx->row_state, ctx->shuffle_count);
}

static int flame_upload(rf_flame* f, float tss_width) {
  rf_ctx* ctx = f->ctx;
  Block b;
  std::string err;
  if (!build_block(f->desc, b, err)) return set_err(ctx, RF_ERR_ARG, err);
  bool animated = false;
  for (int i = 0; i < f->desc.n_xforms; ++i) if (f->desc.xforms[i].rotation_frequency != 0.0f) animated = true;
  if (f->desc.has_final && f->desc.final_xform.rotation_frequency != 0.0f) animated = true;
  const uint32_t T = ctx->sp.temporal_samples;
  f->animated = animated;
  f->n_blocks = animated ? T : 1u;
  f->host_block.assign(1, b);
  f->program = nullptr;  // the structure may have changed
  HIPCHK(ctx, f->blocks.alloc(sizeof(Block) * (animated ? (size_t)T + 1 : 1)));
  const size_t pristine = animated ? T : 0;
  HIPCHK(ctx, hipMemcpyAsync((Block*)f->blocks.p + pristine, &b, sizeof(Block), hipMemcpyHostToDevice, ctx->stream));
  if (animated) {
    float rot[rfd::MAX_XF + 1] = {0};
    for (int i = 0; i < f->desc.n_xforms; ++i) rot[i] = f->desc.xforms[i].rotation_frequency;
    if (f->desc.has_final) rot[f->desc.n_xforms] = f->desc.final_xform.rotation_frequency;
    HIPCHK(ctx, f->rot.alloc(sizeof(rot)));
    HIPCHK(ctx, hipMemcpyAsync(f->rot.p, rot, sizeof(rot), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, rfk::launch_animate((Block*)f->blocks.p, T, tss_width, (const float*)f->rot.p, ctx->stream));
  }
  HIPCHK(ctx, f->palette.alloc(256 * 16));
  HIPCHK(ctx, hipMemcpyAsync(f->palette.p, f->desc.palette, 256 * 16, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  f->needs_update = false;
  // Every warmup starts AUTO's choice afresh: the edited flame's first draw
  // probes the LDS table again and reports a new re-hit fraction.
  f->auto_mode = fixed_point_ok(f->desc) ? RF_HIST_LDS : RF_HIST_QUAD;
  f->auto_probed = false;
  f->last_lds_hit_fraction = -1.0;
  return RF_OK;
}

static int check_desc(rf_ctx* ctx, const rf_flame_desc* d) {
  Block b;
  std::string err;
  if (!build_block(*d, b, err)) return set_err(ctx, RF_ERR_ARG, err);
  return RF_OK;
}

static inline rfk::CompatArgs compat_args(rf_ctx* ctx, rf_flame* f) {
  rfk::CompatArgs a{};
  a.rs = (uint4*)ctx->rs.p;
  a.shuf = (const uint32_t*)ctx->shuf.p;
  a.blocks = (const Block*)f->blocks.p;
  a.block_stride = f->animated ? 1u : 0u;
  a.palette = (const float4*)f->palette.p;
  a.counter = (unsigned long long*)ctx->counter.p;
  a.ppt = (uint32_t)ctx->ppt;
  a.T = ctx->sp.temporal_samples;
  return a;
}

static inline rfk::FastArgs fast_args(rf_ctx* ctx, rf_flame* f) {
  rfk::FastArgs a{};
  a.pos = (float4*)f->local.p;
  a.blocks = (const Block*)f->blocks.p;
  a.block_stride = f->animated ? 1u : 0u;
  a.palette = (const float4*)f->palette.p;
  a.counter = (unsigned long long*)ctx->counter.p;
  a.perms = (const uint16_t*)ctx->perms.p;
  a.seed = ctx->sp.seed;
  a.ppt = (uint32_t)ctx->ppt;
  return a;
}

// Launches the fast engine over the flame's P particles with histogram mode
// hm (1..5): its flame program when programs are on, else the precompiled
// interpreter.  A flame under RF_HIST_AUTO that starts with the LDS table
// gets the sort kernel compiled with it, the mode it usually moves to.
static int fast_launch(rf_ctx* ctx, rf_flame* f, const rfk::FastArgs& a, uint64_t P, int hm) {
  const int W = fast_waves(ctx);
  if (ctx->programs && !f->host_block.empty()) {
    std::vector<std::pair<int, int>> ks{{W, hm}};
    if (ctx->sp.hist_mode == RF_HIST_AUTO && hm == RF_HIST_LDS) ks.push_back({W, RF_HIST_SORT});
    if (ctx->sp.hist_mode == RF_HIST_AUTO && hm == RF_HIST_SORT) ks.push_back({W, RF_HIST_LDS});
    // the flame's cached program when it has the kernels (no structure key
    // to build and look up per launch)
    int dev = 0;
    HIPCHK(ctx, hipGetDevice(&dev));
    const rfj::Program* p = f->program;
    if (!rfj::program_has(p, dev, ks)) {
      std::string err;
      p = rfj::get_program(f->host_block[0], ks, err);
      if (!p) return set_err(ctx, RF_ERR_HIP, err);
      f->program = p;
    }
    f->program_compile_s = rfj::program_compile_seconds(p);
    f->program_used = 1;
    HIPCHK(ctx, rfj::launch_fast(p, a, P, W, hm, ctx->stream));
    return RF_OK;
  }
  f->program_used = 0;
  HIPCHK(ctx, rfk::launch_fast(a, P, W, hm, ctx->stream));
  return RF_OK;
}

static int do_warmup(rf_ctx* ctx, rf_flame* f, uint32_t num_passes, float tss, const uint32_t* rows_in) {
  if (!ctx || !f || f->ctx != ctx) return set_err(ctx, RF_ERR_ARG, "bad context or flame");
  const uint64_t P = ctx->sp.total_particles;
  if (f->generation != ctx->generation) { f->local.release(); f->have_particles = false; f->generation = ctx->generation; }
  HIPCHK(ctx, f->local.alloc(P * 16));
  int rc = flame_upload(f, tss);
  if (rc) return rc;
  if (ctx->sp.engine == RF_ENGINE_COMPAT) {
    if ((rc = ensure_compat(ctx))) return rc;
    std::vector<uint32_t> rows;
    if (rows_in) rows.assign(rows_in, rows_in + 2 * ((size_t)num_passes + 1));
    else next_rows(ctx, num_passes + 1, rows);
    for (auto r : rows) if (r >= ctx->shuffle_count) return set_err(ctx, RF_ERR_ARG, "shuffle row out of range");
    rfk::CompatArgs a = compat_args(ctx, f);
    a.pos_in = (const float4*)ctx->samples.p;
    a.pos_out = (float4*)f->local.p;
    a.random_read = 1; a.random_write = 1; a.first_run = 1; a.do_draw = 0;
    a.in_row = rows[0]; a.out_row = rows[1];
    HIPCHK(ctx, rfk::launch_compat_pass(a, ctx->sp.hist_mode, ctx->stream));
    void* names[2] = {f->local.p, ctx->swap.p};
    a.first_run = 0;
    for (uint32_t i = 0; i < num_passes; ++i) {
      a.pos_in = (const float4*)names[i % 2];
      a.pos_out = (float4*)names[(i + 1) % 2];
      a.in_row = rows[2 * (i + 1)]; a.out_row = rows[2 * (i + 1) + 1];
      HIPCHK(ctx, rfk::launch_compat_pass(a, ctx->sp.hist_mode, ctx->stream));
    }
    if (num_passes % 2) std::swap(f->local.p, ctx->swap.p);
  } else {
    const uint64_t N = 64ull * (uint64_t)fast_waves(ctx);
    if (P % N != 0 || (f->animated && ctx->ppt % N != 0))
      return set_err(ctx, RF_ERR_ARG, "fast engine: particles per temporal sample must be a multiple of 64*block_waves");
    if ((rc = ensure_perms(ctx))) return rc;
    rfk::FastArgs a = fast_args(ctx, f);
    a.launch = ctx->launch++;
    a.n_iter = num_passes + 1;
    a.first_run = 1; a.do_draw = 0;
    a.bins = nullptr;
    // No sample is binned, so any histogram mode's kernel gives the same
    // particles; a flame program uses the kernel its first draw needs (the
    // sort kernel, with its one-workgroup-per-batch grid, stands in for the
    // LDS table's resident grid).
    int hm = ctx->sp.hist_mode;
    if (ctx->programs) {
      hm = ctx->sp.hist_mode == RF_HIST_AUTO ? f->auto_mode : ctx->sp.hist_mode;
      if (hm == RF_HIST_LDS && ctx->sp.hist_mode == RF_HIST_AUTO) hm = RF_HIST_SORT;
    }
    if ((rc = fast_launch(ctx, f, a, P, hm))) return rc;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  f->have_particles = true;
  return RF_OK;
}

// RF_HIST_SORT needs unit sample weights (records carry no weight), palette
// values in [0, 1] (fixed-point tile sums) and at most 2^27 bins.
static bool sort_eligible(const rf_flame* f, uint32_t W, uint32_t H) {
  if ((uint64_t)W * H > ((uint64_t)rfk::SORT_MAX_BUCKETS << rfk::SORT_COARSE_SHIFT)) return false;
  if (!fixed_point_ok(f->desc)) return false;
  for (int i = 0; i < f->desc.n_xforms; ++i)
    if (f->desc.xforms[i].opacity != 1.0f) return false;
  if (f->desc.has_final && f->desc.final_xform.opacity != 1.0f) return false;
  return true;
}

// Passes per sort sub-draw: a draw_to_bins call of more passes runs as
// several chaos launches, each followed by its record passes, that share the
// record buffers; at most 64 passes and 2^28 samples per launch, so the
// buffers stay bounded (about 10 bytes per sample of one sub-draw) and are
// sized once per context instead of growing with every longer draw.
static uint32_t sort_subdraw_passes(uint64_t P) {
  const uint64_t by_samples = P ? ((uint64_t)1 << 28) / P : 1;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, by_samples));
}

// Capacities of the next sort draw from the last one's distribution (20 %
// headroom).  Without one, from the samples the histogram already holds (the
// first draw in AUTO mode is an LDS draw into it), else 3x the mean.  Uploads
// offsets and capacities and grows the record buffers; the plan stays in
// f->sort_meta (layout: boff, bcap, toff, tcap, then the work lists).
static int sort_prepare(rf_ctx* ctx, rf_flame* f, const float4* hist, uint32_t n_bins, uint64_t S,
                        rfk::SortArgs& a) {
  const uint32_t nb = (uint32_t)((n_bins + (1u << rfk::SORT_COARSE_SHIFT) - 1) >> rfk::SORT_COARSE_SHIFT);
  const uint32_t nt = nb * rfk::SORT_TILES_PER_BUCKET;
  const uint32_t nt_used = (uint32_t)((n_bins + (1u << rfk::SORT_TILE_SHIFT) - 1) >> rfk::SORT_TILE_SHIFT);
  // Buffers that only grow, shared by the context's flames.
  auto bind = [&](const uint32_t* dm, size_t fixed, uint32_t n_chunks, uint32_t n_segs) {
    a.cbuf = (uint32_t*)ctx->sort_cbuf.p;
    a.tbuf = (uint32_t*)ctx->sort_tbuf.p;
    a.bcur = (uint32_t*)ctx->sort_bcur.p;
    a.tcur = (uint32_t*)ctx->sort_tcur.p;
    a.boff = dm;
    a.bcap = dm + nb;
    a.toff = dm + 2 * (size_t)nb;
    a.tcap = dm + 2 * (size_t)nb + nt;
    a.chunks = dm + fixed;
    a.segs = dm + fixed + n_chunks;
    a.n_buckets = nb;
    a.n_tiles = nt;
    a.n_chunks = n_chunks;
    a.n_segs = n_segs;
  };
  // A plan covers up to twice the draw it is made for (at most a full
  // sub-draw), so the next draw of up to that size reuses it: the same draw
  // again, or the next of a sequence of doubling draws.  Capacities are upper
  // bounds, a smaller draw leaves part of them empty; a plan for more than
  // four times the draw is replaced, so the record passes do not launch
  // mostly empty workgroups.  Same histogram and no overflow last time: the
  // capacities and work lists on the device still fit.
  if (!f->sort_meta.empty() && S <= f->sort_meta_S && 4 * S >= f->sort_meta_S && f->sort_meta_bins == n_bins &&
      f->sort_bins == n_bins && f->last_sort_overflow == 0.0) {
    bind((const uint32_t*)f->sort_meta_dev.p, f->sort_meta_fixed, f->sort_meta_chunks, f->sort_meta_segs);
    return RF_OK;
  }
  if (f->sort_bins != n_bins || f->sort_bfrac.size() != nb) {
    f->sort_bins = n_bins;
    f->sort_bfrac.assign(nb, 3.0 / nb);
    f->sort_tfrac.assign(nt, 0.0);
    for (uint32_t t = 0; t < nt_used; ++t) f->sort_tfrac[t] = 3.0 / nt_used;
    if (hist) {
      if (ctx->sort_mass.bytes < (size_t)nt_used * 4) HIPCHK(ctx, ctx->sort_mass.alloc((size_t)nt * 4));
      std::vector<float> mass(nt_used);
      HIPCHK(ctx, rfk::launch_sort_tile_mass(hist, n_bins, nt_used, (float*)ctx->sort_mass.p, ctx->stream));
      HIPCHK(ctx, hipMemcpyAsync(mass.data(), ctx->sort_mass.p, (size_t)nt_used * 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      double total = 0.0;
      for (float m : mass) total += m;
      if (total > 0.0) {
        f->sort_bfrac.assign(nb, 0.0);
        for (uint32_t t = 0; t < nt_used; ++t) {
          f->sort_tfrac[t] = mass[t] / total;
          f->sort_bfrac[t / rfk::SORT_TILES_PER_BUCKET] += mass[t] / total;
        }
      }
    }
  }
  const uint64_t S_full = ctx->sp.total_particles * (uint64_t)sort_subdraw_passes(ctx->sp.total_particles);
  const uint64_t Sp = std::max(S, std::min(2 * S, S_full));  // the samples the plan covers
  std::vector<uint32_t> meta(2 * (size_t)nb + 2 * (size_t)nt, 0u);
  uint32_t* boff = meta.data();
  uint32_t* bcap = boff + nb;
  uint32_t* toff = bcap + nb;
  uint32_t* tcap = toff + nt;
  uint64_t cb = 0, ct = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    uint64_t c = (uint64_t)std::ceil((double)Sp * f->sort_bfrac[b] * 1.2) + 2048;
    c = (c + 3) & ~3ull;  // every bucket's records start 16-byte aligned
    boff[b] = (uint32_t)cb;
    bcap[b] = (uint32_t)std::min<uint64_t>(c, 0x7ffffffcull);
    cb += bcap[b];
  }
  for (uint32_t t = 0; t < nt; ++t) {
    uint64_t c = t < nt_used ? (uint64_t)std::ceil((double)Sp * f->sort_tfrac[t] * 1.2) + 512 : 0;
    c = (c + 3) & ~3ull;  // every tile's records start 16-byte aligned
    toff[t] = (uint32_t)ct;
    tcap[t] = (uint32_t)std::min<uint64_t>(c, 0x7ffffffcull);
    ct += tcap[t];
  }
  if (cb >= 0xffffffffull || ct >= 0xffffffffull)
    return set_err(ctx, RF_ERR_LIMIT, "sort histogram: more than 2^32 records per draw; use fewer passes per call");
  // work lists: rebucketing chunks per bucket, summing segments per used tile
  const size_t meta_fixed = meta.size();
  std::vector<uint32_t> jobs;
  for (uint32_t b = 0; b < nb; ++b)
    for (uint32_t k = 0; k * (uint64_t)rfk::SORT_CHUNK < bcap[b]; ++k) {
      if (k > 0xffffu) return set_err(ctx, RF_ERR_LIMIT, "sort histogram: bucket too large");
      jobs.push_back((b << 16) | k);
    }
  const uint32_t n_chunks = (uint32_t)jobs.size();
  // tile segments, the largest first so that the long workgroups start early
  std::vector<std::pair<uint32_t, uint32_t>> segs;  // (records the segment can hold, job)
  for (uint32_t t = 0; t < nt_used; ++t) {
    if (tcap[t] <= (uint32_t)rfk::SORT_SEG16) {
      segs.emplace_back(tcap[t], t << 16);
      continue;
    }
    for (uint32_t k = 0; k * (uint64_t)rfk::SORT_SEG15 < tcap[t]; ++k) {
      if (k > 0xffffu) return set_err(ctx, RF_ERR_LIMIT, "sort histogram: tile too large");
      segs.emplace_back(std::min<uint32_t>(tcap[t] - k * (uint32_t)rfk::SORT_SEG15, rfk::SORT_SEG15), (t << 16) | k);
    }
  }
  order_segments(segs);
  for (const auto& sg : segs) jobs.push_back(sg.second);
  const uint32_t n_segs = (uint32_t)segs.size();
  meta.insert(meta.end(), jobs.begin(), jobs.end());
  // grow to at least what a full sub-draw with measured fractions needs (1.2 S
  // plus the per-bucket and per-tile minimums), so the buffers are allocated
  // once per context rather than at every longer draw
  const uint64_t cb_min = std::min<uint64_t>((uint64_t)(1.25 * (double)S_full) + 2052ull * nb, 0xfffffffcull);
  const uint64_t ct_min = std::min<uint64_t>((uint64_t)(1.25 * (double)S_full) + 516ull * nt_used, 0xfffffffcull);
  if (ctx->sort_cbuf.bytes < cb * 4) HIPCHK(ctx, ctx->sort_cbuf.alloc(std::max(cb, cb_min) * 4));
  if (ctx->sort_tbuf.bytes < ct * 4) HIPCHK(ctx, ctx->sort_tbuf.alloc(std::max(ct, ct_min) * 4));
  if (ctx->sort_bcur.bytes < (size_t)nb * 4) HIPCHK(ctx, ctx->sort_bcur.alloc((size_t)rfk::SORT_MAX_BUCKETS * 4));
  if (ctx->sort_tcur.bytes < (size_t)nt * 4)
    HIPCHK(ctx, ctx->sort_tcur.alloc((size_t)rfk::SORT_MAX_BUCKETS * rfk::SORT_TILES_PER_BUCKET * 4));
  if (f->sort_meta_dev.bytes < meta.size() * 4) HIPCHK(ctx, f->sort_meta_dev.alloc(meta.size() * 4 * 2));
  // every draw ends with a stream synchronisation, so the stage is free here;
  // sized once for any plan up to 1 MiB (pinned allocations are slow, and a
  // sequence of growing draws would otherwise grow it step by step)
  HIPCHK(ctx, ctx->meta_stage.alloc(std::max<size_t>(meta.size() * 4, (size_t)1 << 20)));
  memcpy(ctx->meta_stage.p, meta.data(), meta.size() * 4);
  HIPCHK(ctx, hipMemcpyAsync(f->sort_meta_dev.p, ctx->meta_stage.p, meta.size() * 4, hipMemcpyHostToDevice,
                             ctx->stream));
  bind((const uint32_t*)f->sort_meta_dev.p, meta_fixed, n_chunks, n_segs);
  f->sort_meta.swap(meta);
  f->sort_meta_S = Sp;
  f->sort_meta_bins = n_bins;
  f->sort_meta_fixed = meta_fixed;
  f->sort_meta_chunks = n_chunks;
  f->sort_meta_segs = n_segs;
  return RF_OK;
}

// After a sort draw: record where the samples went and how many overflowed.
// bc, tc: the draw's bucket and tile cursors, already copied to the host.
// S and drawn: the samples attempted and binned by the draw whose cursors
// these are (the last sub-draw of a long call).
static void sort_finish(rf_flame* f, const rfk::SortArgs& a, uint64_t S, uint64_t drawn,
                        const std::vector<uint32_t>& meta, const uint32_t* bc, const uint32_t* tc) {
  const uint32_t* bcap = meta.data() + a.n_buckets;
  const uint32_t* tcap = meta.data() + 2 * (size_t)a.n_buckets + a.n_tiles;
  uint64_t ovf = 0;
  const double inv = S ? 1.0 / (double)S : 0.0;
  for (uint32_t b = 0; b < a.n_buckets; ++b) {
    if (bc[b] > bcap[b]) ovf += bc[b] - bcap[b];
    f->sort_bfrac[b] = bc[b] * inv;
  }
  for (uint32_t t = 0; t < a.n_tiles; ++t) {
    if (tc[t] > tcap[t]) ovf += tc[t] - tcap[t];
    // a bucket's overflow never reaches its tiles: keep their share at least the mean
    f->sort_tfrac[t] = tc[t] * inv;
  }
  f->last_sort_overflow = drawn ? (double)ovf / (double)drawn : 0.0;
}

// RF_HIST_AUTO keeps the LDS bin cache while at least this fraction of the
// binned samples finds its bin in a workgroup's table.  Measured on MI355X the
// cache already breaks even with the quad path at a 4 % hit fraction (a
// 7680x4320 view of an 8-xform flame), so it is only dropped when the table
// is nearly useless.
static constexpr double kAutoLdsMinHits = 0.02;
// ... and prefers the sort path below this fraction when the flame allows it:
// only a table that catches nearly every sample (a few hundred hot bins, where
// the sort path's tiles would be hot too) beats it.
static constexpr double kAutoSortMaxHits = 0.90;

static int do_draw(rf_ctx* ctx, rf_flame* f, rf_hist* h, uint32_t num_iter, const uint32_t* rows_in,
                   uint64_t* drawn) {
  // AUTO mode's first draw of a flame measures the LDS table's re-hits on one
  // pass and runs the other passes in the mode that measurement picks (a
  // whole 7680x4320 draw through the LDS table costs five sort draws).
  if (ctx && f && h && ctx->sp.engine == RF_ENGINE_FAST && ctx->sp.hist_mode == RF_HIST_AUTO &&
      f->auto_mode == RF_HIST_LDS && !f->auto_probed && num_iter > 1 && !rows_in) {
    f->auto_probed = true;
    uint64_t d1 = 0, d2 = 0;
    int rc = do_draw(ctx, f, h, 1, nullptr, &d1);
    if (rc) return rc;
    const float ms1 = ctx->last_draw_ms;
    rc = do_draw(ctx, f, h, num_iter - 1, nullptr, &d2);
    if (rc) return rc;
    ctx->last_draw_ms += ms1;
    if (drawn) *drawn = d1 + d2;
    return RF_OK;
  }
  if (!ctx || !f || !h || f->ctx != ctx || h->ctx != ctx) return set_err(ctx, RF_ERR_ARG, "bad context, flame or histogram");
  if (!f->have_particles || f->generation != ctx->generation || f->needs_update)
    return set_err(ctx, RF_ERR_STATE, "flame needs a warmup before draw_to_bins");
  const uint64_t P = ctx->sp.total_particles;
  float ss[6];
  screen_affine(f->desc, h->w, h->h, ss);
  constexpr size_t counter_bytes = 8ull * rfk::COUNTER_SLOTS * rfk::COUNTER_STRIDE;
  constexpr uint32_t counter_words = (uint32_t)(counter_bytes / 8);
  // sort draws clear the counters in their first cursor reset
  bool counter_cleared = false;
  const int hm = ctx->sp.hist_mode == RF_HIST_AUTO ? f->auto_mode : ctx->sp.hist_mode;
  bool sort_used = false;
  rfk::SortArgs sort_args{};
  uint64_t sort_S = 0;
  if (!ctx->ev[2]) { HIPCHK(ctx, hipEventCreate(&ctx->ev[2])); HIPCHK(ctx, hipEventCreate(&ctx->ev[3])); }
  if (ctx->sp.engine == RF_ENGINE_COMPAT) {
    int rc = ensure_compat(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipMemsetAsync(ctx->counter.p, 0, counter_bytes, ctx->stream));
    counter_cleared = true;
    std::vector<uint32_t> rows;
    if (rows_in) rows.assign(rows_in, rows_in + 2 * (size_t)num_iter);
    else next_rows(ctx, num_iter, rows);
    for (auto r : rows) if (r >= ctx->shuffle_count) return set_err(ctx, RF_ERR_ARG, "shuffle row out of range");
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    rfk::CompatArgs a = compat_args(ctx, f);
    a.random_read = 1; a.random_write = 0; a.first_run = 0; a.do_draw = 1;
    a.bins = (float*)h->ptr;
    a.W = h->w; a.H = h->h;
    memcpy(a.ss, ss, sizeof(ss));
    void* names[2] = {f->local.p, ctx->swap.p};
    for (uint32_t i = 0; i < num_iter; ++i) {
      a.pos_in = (const float4*)names[i % 2];
      a.pos_out = (float4*)names[(i + 1) % 2];
      a.in_row = rows[2 * i]; a.out_row = rows[2 * i + 1];
      HIPCHK(ctx, rfk::launch_compat_pass(a, ctx->sp.hist_mode, ctx->stream));
    }
    if (num_iter % 2) std::swap(f->local.p, ctx->swap.p);
  } else {
    int rc = ensure_perms(ctx);
    if (rc) return rc;
    rfk::FastArgs a = fast_args(ctx, f);
    a.launch = ctx->launch++;
    a.n_iter = num_iter;
    a.first_run = 0; a.do_draw = 1;
    a.bins = (float*)h->ptr;
    a.W = h->w; a.H = h->h;
    memcpy(a.ss, ss, sizeof(ss));
    int hm_run = hm;
    if (hm_run == RF_HIST_SORT && !sort_eligible(f, h->w, h->h)) hm_run = RF_HIST_QUAD;
    if (hm_run == RF_HIST_LDS && !fixed_point_ok(f->desc)) hm_run = RF_HIST_QUAD;
    // the sort path runs long calls as sub-draws of at most `sub` passes,
    // all planned with the capacities of a full sub-draw
    const uint32_t sub = hm_run == RF_HIST_SORT ? std::min(num_iter, sort_subdraw_passes(P)) : num_iter;
    if (hm_run == RF_HIST_SORT && num_iter) {
      const int rc2 = sort_prepare(ctx, f, (const float4*)h->ptr, h->w * h->h, P * (uint64_t)sub, a.sort);
      if (rc2) return rc2;
    }
    if (hm_run != RF_HIST_SORT || !num_iter) {
      HIPCHK(ctx, hipMemsetAsync(ctx->counter.p, 0, counter_bytes, ctx->stream));
      counter_cleared = true;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    uint32_t last = num_iter;
    for (uint32_t done = 0; done < num_iter; done += last) {
      last = std::min(sub, num_iter - done);
      if (done) a.launch = ctx->launch++;
      a.n_iter = last;
      if (hm_run == RF_HIST_SORT) {
        HIPCHK(ctx, rfk::launch_sort_reset(a.sort, ctx->stream, (unsigned long long*)ctx->counter.p,
                                           counter_cleared ? 0u : counter_words));
        counter_cleared = true;
      }
      const int rc3 = fast_launch(ctx, f, a, P, hm_run);
      if (rc3) return rc3;
      if (hm_run == RF_HIST_SORT) {
        HIPCHK(ctx, rfk::launch_sort_rebucket(a.sort, a.bins, a.palette, ctx->stream));
        HIPCHK(ctx, rfk::launch_sort_tiles(a.sort, a.bins, h->w * h->h, a.palette, ctx->stream));
      }
    }
    sort_used = hm_run == RF_HIST_SORT && num_iter;
    sort_args = a.sort;
    sort_S = P * (uint64_t)last;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  // counter slots, then the sort cursors, read back by one kernel into pinned
  // memory behind one synchronisation
  const size_t cur_words = sort_used ? (size_t)sort_args.n_buckets + sort_args.n_tiles : 0;
  // sized once for the largest cursor arrays
  HIPCHK(ctx, ctx->readback.alloc(std::max<size_t>(
      counter_bytes + cur_words * 4,
      counter_bytes + 4ull * rfk::SORT_MAX_BUCKETS * (1 + rfk::SORT_TILES_PER_BUCKET))));
  const uint64_t* slots = static_cast<const uint64_t*>(ctx->readback.p);
  uint32_t* cur_host = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->readback.p) + counter_bytes);
  HIPCHK(ctx, rfk::launch_draw_readback((const unsigned long long*)ctx->counter.p, counter_words,
                                        sort_used ? sort_args.bcur : nullptr, sort_args.n_buckets,
                                        sort_used ? sort_args.tcur : nullptr, sort_args.n_tiles,
                                        (uint32_t*)ctx->readback.dev, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t cnt[2] = {0, 0};
  for (int k = 0; k < rfk::COUNTER_SLOTS; ++k) {
    cnt[0] += slots[k * rfk::COUNTER_STRIDE];
    cnt[1] += slots[k * rfk::COUNTER_STRIDE + 1];
  }
  if (sort_used) {
    // the cursors are the last sub-draw's: its share of the binned samples
    const uint64_t S_all = P * (uint64_t)num_iter;
    const uint64_t d_last = S_all ? (uint64_t)((double)cnt[0] * (double)sort_S / (double)S_all) : 0;
    sort_finish(f, sort_args, sort_S, d_last, f->sort_meta, cur_host, cur_host + sort_args.n_buckets);
  }
  if (ctx->sp.engine == RF_ENGINE_FAST && hm == RF_HIST_LDS && cnt[0]) {
    f->last_lds_hit_fraction = (double)cnt[1] / (double)cnt[0];
    if (ctx->sp.hist_mode == RF_HIST_AUTO && f->last_lds_hit_fraction < kAutoSortMaxHits) {
      if (sort_eligible(f, h->w, h->h)) f->auto_mode = RF_HIST_SORT;
      else if (f->last_lds_hit_fraction < kAutoLdsMinHits) f->auto_mode = RF_HIST_QUAD;
    }
  }
  float ms = 0.0f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  ctx->last_draw_ms = ms;
  if (drawn) *drawn = cnt[0];
  return RF_OK;
}

// DE arguments for histogram h; grows the histogram's radius buffer and the
// context's weight table to what the launch needs.
static int de_args(rf_ctx* ctx, rf_hist* h, const rf_de_params* de, rfk::DEArgs& a) {
  a = rfk::DEArgs{};
  a.hist = (const float4*)h->ptr;
  a.W = h->w; a.H = h->h;
  if (de) {
    a.R = std::max(0, std::min(de->estimator_radius, 100));  // src/main.cpp:502
    a.rmin = std::max(0, std::min(de->estimator_min, 100));
    a.curve = de->estimator_curve;
  }
  const size_t nb = (size_t)h->w * h->h, wb = rfk::de_weight_bytes(std::max(a.R, a.rmin));
  if (h->de_rad.bytes < nb) HIPCHK(ctx, h->de_rad.alloc(nb));
  if (ctx->de_wt.bytes < wb) HIPCHK(ctx, ctx->de_wt.alloc(wb));
  a.rad = (signed char*)h->de_rad.p;
  a.wt = (float*)ctx->de_wt.p;
  return RF_OK;
}

// ------------------------------------------------ allocation-free draws
// rf_draw_to_bins_noalloc keeps its draw counters (the COUNTER_SLOTS x
// COUNTER_STRIDE u64 block the chaos-game kernels add their binned and
// LDS-hit counts to) in this file-scope device array instead of the context's
// buffer.  Every GPU holds its own copy for the life of the process.  In host
// code `scratch` is only a stand-in for the symbol: kernels get the address
// hipGetSymbolAddress returned for the device (g_scratch_ptr), which
// scratch_init takes once per device with that device current.
constexpr size_t kScratchWords = (size_t)rfk::COUNTER_SLOTS * rfk::COUNTER_STRIDE;
static_assert(kScratchWords * 8 <= RF_NOALLOC_SCRATCH_BYTES, "draw counters exceed the no-alloc scratch array");
static __device__ __align__(256) unsigned long long scratch[RF_NOALLOC_SCRATCH_BYTES / 8];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned long long* g_scratch_ptr[kScratchDevices] = {nullptr};

// Takes the address of `scratch` on `dev`, which must be the current device.
static hipError_t scratch_init(int dev) {
  if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
  return hipSuccess;
}

static unsigned long long* scratch_ptr(int dev) {
  if (dev < 0 || dev >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

// Reason of the last failed rf_create (rf_last_error(NULL)).
static thread_local std::string g_create_err;

// ================================================================== C ABI
extern "C" {

int rf_variation_count(void) { return RF_GEN_NUM_VARIATIONS; }
const char* rf_variation_name(int id) { const rf_var_info* v = var_info(id); return v ? v->name : nullptr; }
int rf_variation_id(const char* name) {
  if (!name) return -1;
  auto& idx = pindex();
  auto it = idx.vars.find(name);
  return it == idx.vars.end() ? -1 : it->second;
}
int rf_variation_num_params(int id) { const rf_var_info* v = var_info(id); return v ? v->n_params : -1; }
const char* rf_variation_param_name(int id, int k) {
  const rf_var_info* v = var_info(id);
  return (v && k >= 0 && k < v->n_params) ? v->params[k] : nullptr;
}
int rf_variation_flags(int id) {
  const rf_var_info* v = var_info(id);
  return v ? (int)(v->flags | (v->uses_rand ? 0x100u : 0u)) : -1;
}

const char* rf_last_error(const rf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }
double rf_last_draw_ms(const rf_ctx* ctx) { return ctx ? ctx->last_draw_ms : -1.0; }
void* rf_stream(rf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rf_create(const rf_sim_params* p, rf_ctx** out) {
  if (!p || !out) return RF_ERR_ARG;
  *out = nullptr;
  rf_ctx* ctx = new rf_ctx();
  ctx->sp = *p;
  if (ctx->sp.shuffle_count == 0) ctx->sp.shuffle_count = 1024;
  if (const char* env = getenv("RF_FLAME_PROGRAMS")) ctx->programs = strcmp(env, "0") != 0;
  if (p->block_waves != 0 && p->block_waves != 4 && p->block_waves != 8) {
    g_create_err = "block_waves must be 0 (default 4), 4 or 8";
    delete ctx;
    return RF_ERR_ARG;
  }
  if (p->engine != RF_ENGINE_COMPAT && p->engine != RF_ENGINE_FAST) {
    g_create_err = "unknown engine";
    delete ctx;
    return RF_ERR_ARG;
  }
  if (p->hist_mode < RF_HIST_AUTO || p->hist_mode > RF_HIST_SORT) {
    g_create_err = "unknown hist mode";
    delete ctx;
    return RF_ERR_ARG;
  }
  int rc = validate_sim(ctx, p->total_particles, p->temporal_samples);
  if (rc) { g_create_err = ctx->err; delete ctx; return rc; }
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) { g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e); delete ctx; return RF_ERR_HIP; }
  // A blocking stream: it orders with the legacy default stream, so device
  // buffers that a caller (e.g. torch) filled there are ready before our
  // kernels read or write them, and the caller's later work sees our results.
  e = hipStreamCreate(&ctx->stream);
  if (e != hipSuccess) { g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e); delete ctx; return RF_ERR_HIP; }
  e = ctx->counter.alloc(8ull * rfk::COUNTER_SLOTS * rfk::COUNTER_STRIDE);
  if (e != hipSuccess) { g_create_err = std::string("hipMalloc: ") + hipGetErrorString(e); rf_destroy(ctx); return RF_ERR_HIP; }
  e = scratch_init(p->device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    rf_destroy(ctx);
    return RF_ERR_HIP;
  }
  ctx->ppt = p->total_particles / p->temporal_samples;
  ctx->shuffle_count = ctx->sp.shuffle_count;
  ctx->row_state = p->seed ^ 0x0123456789ABCDEFull;
  *out = ctx;
  return RF_OK;
}

void rf_destroy(rf_ctx* ctx) {
  if (!ctx) return;
  if (ctx->render_hist) rf_hist_destroy(ctx->render_hist);
  if (ctx->stream) { (void)hipStreamSynchronize(ctx->stream); (void)hipStreamDestroy(ctx->stream); }
  for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
  delete ctx;
}

int rf_set_sim_parameters(rf_ctx* ctx, uint64_t P, uint32_t T, uint32_t R) {
  if (!ctx) return RF_ERR_ARG;
  int rc = validate_sim(ctx, P, T);
  if (rc) return rc;
  ctx->sp.total_particles = P;
  ctx->sp.temporal_samples = T;
  ctx->sp.shuffle_count = R ? R : 1024;
  ctx->shuffle_count = ctx->sp.shuffle_count;
  ctx->ppt = P / T;
  ctx->swap.release(); ctx->rs.release(); ctx->shuf.release(); ctx->samples.release();
  ctx->compat_ready = false;
  ctx->generation++;  // invalidates every flame's particle state (src/flame.cpp:153-157)
  return RF_OK;
}

int rf_set_engine(rf_ctx* ctx, int32_t engine, int32_t hist_mode) {
  if (!ctx) return RF_ERR_ARG;
  if (engine != RF_ENGINE_COMPAT && engine != RF_ENGINE_FAST) return set_err(ctx, RF_ERR_ARG, "unknown engine");
  if (hist_mode < RF_HIST_AUTO || hist_mode > RF_HIST_SORT) return set_err(ctx, RF_ERR_ARG, "unknown hist mode");
  if (engine != ctx->sp.engine) ctx->generation++;  // particle layouts differ between engines
  ctx->sp.engine = engine;
  ctx->sp.hist_mode = hist_mode;
  return RF_OK;
}

int rf_flame_parse_desc(const char* xml, size_t len, rf_flame_desc* out, char* err, size_t err_len) {
  if (!xml || !out) return RF_ERR_ARG;
  std::string e;
  int rc = parse_flam3(xml, len, out, e);
  if (rc && err && err_len) { snprintf(err, err_len, "%s", e.c_str()); }
  return rc;
}

int rf_flame_create(rf_ctx* ctx, const rf_flame_desc* desc, rf_flame** out) {
  if (!ctx || !desc || !out) return set_err(ctx, RF_ERR_ARG, "null argument");
  *out = nullptr;
  int rc = check_desc(ctx, desc);
  if (rc) return rc;
  rf_flame* f = new rf_flame();
  f->ctx = ctx;
  f->desc = *desc;
  f->generation = ctx->generation;
  *out = f;
  return RF_OK;
}

int rf_flame_parse(rf_ctx* ctx, const char* xml, size_t len, rf_flame** out) {
  if (!ctx || !xml || !out) return set_err(ctx, RF_ERR_ARG, "null argument");
  rf_flame_desc* d = new rf_flame_desc();
  std::string e;
  int rc = parse_flam3(xml, len, d, e);
  if (rc) { delete d; return set_err(ctx, rc, e); }
  rc = rf_flame_create(ctx, d, out);
  delete d;
  return rc;
}

int rf_flame_load(rf_ctx* ctx, const char* path, rf_flame** out) {
  if (!ctx || !path || !out) return set_err(ctx, RF_ERR_ARG, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return set_err(ctx, RF_ERR_IO, std::string("cannot open ") + path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  return rf_flame_parse(ctx, s.data(), s.size(), out);
}

int rf_flame_update(rf_flame* f, const rf_flame_desc* desc) {
  if (!f || !desc) return RF_ERR_ARG;
  int rc = check_desc(f->ctx, desc);
  if (rc) return rc;
  f->desc = *desc;
  f->needs_update = true;
  return RF_OK;
}

int rf_flame_get_desc(const rf_flame* f, rf_flame_desc* out) {
  if (!f || !out) return RF_ERR_ARG;
  *out = f->desc;
  return RF_OK;
}

void rf_flame_destroy(rf_flame* f) { delete f; }

int rf_flame_param_block(const rf_flame_desc* d, float* out, int cap) {
  if (!d) return RF_ERR_ARG;
  std::vector<float> v = reference_param_block(*d);
  if (out) for (int i = 0; i < cap && i < (int)v.size(); ++i) out[i] = v[i];
  return (int)v.size();
}

int rf_flame_validate(const rf_flame_desc* d, char* err, size_t err_len) {
  if (!d) return RF_ERR_ARG;
  Block* b = new Block();
  std::string e;
  bool ok = build_block(*d, *b, e);
  delete b;
  if (!ok && err && err_len) snprintf(err, err_len, "%s", e.c_str());
  return ok ? RF_OK : RF_ERR_ARG;
}

int rf_rgba8_device(rf_ctx* ctx, const void* in, void* out, uint32_t w, uint32_t h) {
  if (!ctx) return RF_ERR_ARG;
  if (!in || !out || w == 0 || h == 0) return set_err(ctx, RF_ERR_ARG, "rgba8: null pointer or empty image");
  const uint64_t n = (uint64_t)w * h;
  if (n > 0xffffffffull) return set_err(ctx, RF_ERR_LIMIT, "rgba8: more than 2^32 pixels");
  HIPCHK(ctx, rfk::launch_rgba8((const float4*)in, (uchar4*)out, (uint32_t)n, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

double rf_flame_sort_overflow(const rf_flame* f) { return f ? f->last_sort_overflow : -1.0; }

int rf_set_flame_programs(rf_ctx* ctx, int32_t enable) {
  if (!ctx) return RF_ERR_ARG;
  if (enable != 0 && enable != 1) return set_err(ctx, RF_ERR_ARG, "enable must be 0 or 1");
  ctx->programs = enable == 1;
  return RF_OK;
}

int rf_flame_program_info(const rf_flame* f, double* compile_seconds) {
  if (!f) return RF_ERR_ARG;
  if (compile_seconds) *compile_seconds = f->program_compile_s;
  return f->program_used;
}

static int program_request(const rf_flame_desc* d, int32_t waves, int32_t hm, Block& b, std::string& err) {
  if (!d) { err = "null description"; return RF_ERR_ARG; }
  if (waves != 0 && waves != 4 && waves != 8) { err = "block_waves must be 0, 4 or 8"; return RF_ERR_ARG; }
  const bool batch = hm == rfj::kBatchKernel + RF_HIST_QUAD || hm == rfj::kBatchKernel + RF_HIST_LDS ||
                     hm == rfj::kBatchKernel + RF_HIST_SORT;
  if (!batch && (hm < RF_HIST_QUAD || hm > RF_HIST_SORT)) {
    err = "hist_mode must be 1..5, or 16 + 1, 16 + 4 or 16 + 5 for the batched renderer";
    return RF_ERR_ARG;
  }
  if (!build_block(*d, b, err)) return RF_ERR_ARG;
  return RF_OK;
}

int rf_flame_program_source(const rf_flame_desc* d, int32_t waves, int32_t hm, char* buf, size_t cap) {
  std::unique_ptr<Block> b(new Block);
  std::string err;
  if (program_request(d, waves, hm, *b, err)) return RF_ERR_ARG;
  const std::string src = rfj::program_source(*b, {{waves ? waves : 8, hm}});
  if (buf && cap) {
    const size_t n = std::min(cap - 1, src.size());
    memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return (int)src.size() + 1;
}

int rf_flame_program_build(const rf_flame_desc* d, int32_t waves, int32_t hm, const char* arch, void* code,
                           size_t cap, size_t* size, char* log, size_t log_len, double* seconds) {
  std::unique_ptr<Block> b(new Block);
  std::string err;
  auto put_log = [&](const std::string& m) {
    if (log && log_len) {
      const size_t n = std::min(log_len - 1, m.size());
      memcpy(log, m.data(), n);
      log[n] = '\0';
    }
  };
  if (program_request(d, waves, hm, *b, err)) { put_log(err); return RF_ERR_ARG; }
  if (!arch || !*arch) { put_log("arch is empty"); return RF_ERR_ARG; }
  std::vector<char> obj;
  std::string clog;
  if (!rfj::compile(rfj::program_source(*b, {{waves ? waves : 8, hm}}), arch, obj, clog, seconds)) {
    put_log(clog);
    return RF_ERR_HIP;
  }
  put_log(clog);
  if (size) *size = obj.size();
  if (code && cap >= obj.size()) memcpy(code, obj.data(), obj.size());
  return (code && cap < obj.size()) ? RF_ERR_ARG : RF_OK;
}

int rf_flame_hist_mode(const rf_flame* f, double* frac) {
  if (!f) return RF_ERR_ARG;
  if (frac) *frac = f->last_lds_hit_fraction;
  const int m = f->ctx->sp.hist_mode;
  return m == RF_HIST_AUTO ? f->auto_mode : m;
}

int rf_screen_affine(const rf_flame_desc* d, uint32_t w, uint32_t h, float out[6]) {
  if (!d || !out) return RF_ERR_ARG;
  screen_affine(*d, w, h, out);
  return RF_OK;
}

int rf_hist_create(rf_ctx* ctx, uint32_t w, uint32_t h, rf_hist** out) {
  if (!ctx || !out || w == 0 || h == 0) return set_err(ctx, RF_ERR_ARG, "bad histogram size");
  if ((uint64_t)w * h * 4 > 0xffffffffull) return set_err(ctx, RF_ERR_LIMIT, "histogram larger than 2^30 bins");
  rf_hist* hs = new rf_hist();
  hs->ctx = ctx; hs->w = w; hs->h = h; hs->owned = true;
  hipError_t e = hs->own.alloc((size_t)w * h * 16);
  if (e != hipSuccess) { delete hs; return set_err(ctx, RF_ERR_HIP, std::string("hist alloc: ") + hipGetErrorString(e)); }
  hs->ptr = hs->own.p;
  e = hipMemsetAsync(hs->ptr, 0, (size_t)w * h * 16, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) { delete hs; return set_err(ctx, RF_ERR_HIP, hipGetErrorString(e)); }
  *out = hs;
  return RF_OK;
}

int rf_hist_wrap(rf_ctx* ctx, void* dev, uint32_t w, uint32_t h, rf_hist** out) {
  if (!ctx || !dev || !out || w == 0 || h == 0) return set_err(ctx, RF_ERR_ARG, "bad histogram");
  if ((uint64_t)w * h * 4 > 0xffffffffull) return set_err(ctx, RF_ERR_LIMIT, "histogram larger than 2^30 bins");
  rf_hist* hs = new rf_hist();
  hs->ctx = ctx; hs->w = w; hs->h = h; hs->ptr = dev; hs->owned = false;
  *out = hs;
  return RF_OK;
}

void rf_hist_destroy(rf_hist* h) { delete h; }

int rf_hist_zero(rf_hist* h) {
  if (!h) return RF_ERR_ARG;
  HIPCHK(h->ctx, hipMemsetAsync(h->ptr, 0, (size_t)h->w * h->h * 16, h->ctx->stream));
  HIPCHK(h->ctx, hipStreamSynchronize(h->ctx->stream));
  return RF_OK;
}

int rf_hist_download(rf_hist* h, float* host) {
  if (!h || !host) return RF_ERR_ARG;
  HIPCHK(h->ctx, hipMemcpyAsync(host, h->ptr, (size_t)h->w * h->h * 16, hipMemcpyDeviceToHost, h->ctx->stream));
  HIPCHK(h->ctx, hipStreamSynchronize(h->ctx->stream));
  return RF_OK;
}

int rf_hist_upload(rf_hist* h, const float* host) {
  if (!h || !host) return RF_ERR_ARG;
  HIPCHK(h->ctx, hipMemcpyAsync(h->ptr, host, (size_t)h->w * h->h * 16, hipMemcpyHostToDevice, h->ctx->stream));
  HIPCHK(h->ctx, hipStreamSynchronize(h->ctx->stream));
  return RF_OK;
}

void* rf_hist_device_ptr(rf_hist* h) { return h ? h->ptr : nullptr; }

int rf_warmup(rf_ctx* ctx, rf_flame* f, uint32_t n, float tss) { return do_warmup(ctx, f, n, tss, nullptr); }
int rf_warmup_rows(rf_ctx* ctx, rf_flame* f, uint32_t n, float tss, const uint32_t* rows) {
  if (!rows) return set_err(ctx, RF_ERR_ARG, "rows is null");
  return do_warmup(ctx, f, n, tss, rows);
}
int rf_draw_to_bins(rf_ctx* ctx, rf_flame* f, rf_hist* h, uint32_t n, uint64_t* drawn) {
  return do_draw(ctx, f, h, n, nullptr, drawn);
}
int rf_draw_to_bins_rows(rf_ctx* ctx, rf_flame* f, rf_hist* h, uint32_t n, const uint32_t* rows, uint64_t* drawn) {
  if (!rows) return set_err(ctx, RF_ERR_ARG, "rows is null");
  return do_draw(ctx, f, h, n, rows, drawn);
}

int rf_draw_to_bins_noalloc(rf_ctx* ctx, rf_flame* f, rf_hist* h, uint32_t num_iter, uint64_t* drawn) {
  if (!ctx || !f || !h || f->ctx != ctx || h->ctx != ctx) return set_err(ctx, RF_ERR_ARG, "bad context, flame or histogram");
  if (ctx->sp.engine != RF_ENGINE_FAST) return set_err(ctx, RF_ERR_ARG, "rf_draw_to_bins_noalloc needs the fast engine");
  if (!f->have_particles || f->generation != ctx->generation || f->needs_update)
    return set_err(ctx, RF_ERR_STATE, "flame needs a warmup before draw_to_bins");
  if (ctx->perm_waves != fast_waves(ctx) || !ctx->perms.p)
    return set_err(ctx, RF_ERR_STATE, "flame needs a warmup before draw_to_bins");
  unsigned long long* cnt_dev = scratch_ptr(ctx->sp.device);
  if (!cnt_dev) return set_err(ctx, RF_ERR_STATE, "no scratch address for this device");
  if (kScratchWords * 8 > RF_NOALLOC_SCRATCH_BYTES)
    return set_err(ctx, RF_ERR_LIMIT, "draw counters exceed RF_NOALLOC_SCRATCH_BYTES");
  // The direct histogram modes only: the sort path needs record buffers the
  // size of the draw.  AUTO runs its current choice, the LDS table in place
  // of the sort path, and never probes.
  int hm = ctx->sp.hist_mode == RF_HIST_AUTO ? f->auto_mode : ctx->sp.hist_mode;
  if (hm == RF_HIST_SORT) hm = RF_HIST_LDS;
  if (hm == RF_HIST_LDS && !fixed_point_ok(f->desc)) hm = RF_HIST_QUAD;
  uint64_t host[kScratchWords];
  HIPCHK(ctx, rfk::launch_zero_u64(cnt_dev, (uint32_t)kScratchWords, ctx->stream));
  if (num_iter) {
    const uint64_t P = ctx->sp.total_particles;
    rfk::FastArgs a = fast_args(ctx, f);
    a.counter = cnt_dev;
    a.launch = ctx->launch++;
    a.n_iter = num_iter;
    a.first_run = 0; a.do_draw = 1;
    a.bins = (float*)h->ptr;
    a.W = h->w; a.H = h->h;
    screen_affine(f->desc, h->w, h->h, a.ss);
    const int W = fast_waves(ctx);
    // a compiled flame program when one is already loaded (compiling would
    // allocate), else the precompiled interpreter: both give the same bits
    const rfj::Program* prog = (ctx->programs && !f->host_block.empty()) ?
                               rfj::find_program(f->host_block[0], {W, hm}) : nullptr;
    if (prog) HIPCHK(ctx, rfj::launch_fast(prog, a, P, W, hm, ctx->stream));
    else HIPCHK(ctx, rfk::launch_fast(a, P, W, hm, ctx->stream));
  }
  HIPCHK(ctx, hipMemcpyAsync(host, cnt_dev, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t n = 0, lds_hits = 0;
  for (int k = 0; k < rfk::COUNTER_SLOTS; ++k) {
    n += host[k * rfk::COUNTER_STRIDE];
    lds_hits += host[k * rfk::COUNTER_STRIDE + 1];
  }
  if (hm == RF_HIST_LDS && n) f->last_lds_hit_fraction = (double)lds_hits / (double)n;
  ctx->last_draw_ms = -1.0;  // not timed: events are created on first use
  if (drawn) *drawn = n;
  return RF_OK;
}

int rf_density_estimate(rf_ctx* ctx, rf_hist* h, const rf_de_params* de, void* out) {
  if (!ctx || !h || !de) return set_err(ctx, RF_ERR_ARG, "null argument");
  if (!out) { HIPCHK(ctx, h->image.alloc((size_t)h->w * h->h * 16)); out = h->image.p; }
  rfk::DEArgs a;
  int rc = de_args(ctx, h, de, a);
  if (rc) return rc;
  a.de_out = (float4*)out;
  a.do_tonemap = 0;
  HIPCHK(ctx, rfk::launch_de_tonemap(a, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_tonemap(rf_ctx* ctx, const void* in, void* out, uint32_t w, uint32_t h, const rf_tonemap_params* tm) {
  if (!ctx || !in || !out || !tm) return set_err(ctx, RF_ERR_ARG, "null argument");
  HIPCHK(ctx, rfk::launch_tonemap((const float4*)in, (float4*)out, w * h, tm->gamma, tm->brightness, tm->vibrancy,
                                  tm->scale_constant, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_density_tonemap(rf_ctx* ctx, rf_hist* h, const rf_de_params* de, const rf_tonemap_params* tm, void* out) {
  if (!ctx || !h || !tm) return set_err(ctx, RF_ERR_ARG, "null argument");
  if (!out) { HIPCHK(ctx, h->image.alloc((size_t)h->w * h->h * 16)); out = h->image.p; }
  rfk::DEArgs a;
  int rc = de_args(ctx, h, de, a);
  if (rc) return rc;
  a.img_out = (float4*)out;
  a.do_tonemap = 1;
  a.gamma = tm->gamma; a.brightness = tm->brightness; a.vibrancy = tm->vibrancy; a.k = tm->scale_constant;
  HIPCHK(ctx, rfk::launch_de_tonemap(a, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RF_OK;
}

int rf_render(rf_ctx* ctx, rf_flame* f, const rf_render_params* rp, float* hist_out, float* image_out, rf_stats* st) {
  if (!ctx || !f || !rp) return set_err(ctx, RF_ERR_ARG, "null argument");
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  if (!ctx->render_hist || ctx->render_hist->w != rp->width || ctx->render_hist->h != rp->height) {
    if (ctx->render_hist) { rf_hist_destroy(ctx->render_hist); ctx->render_hist = nullptr; }
    int rc = rf_hist_create(ctx, rp->width, rp->height, &ctx->render_hist);
    if (rc) return rc;
  }
  rf_hist* h = ctx->render_hist;
  int rc = rf_hist_zero(h);
  if (rc) return rc;
  rf_stats s{};
  auto t0 = clk::now();
  if ((rc = rf_warmup(ctx, f, rp->warmup_passes, rp->tss_width))) return rc;
  auto t1 = clk::now();
  for (uint32_t k = 0; k < rp->draw_calls; ++k) {
    uint64_t d = 0;
    if ((rc = rf_draw_to_bins(ctx, f, h, rp->draw_passes, &d))) return rc;
    s.drawn += d;
  }
  auto t2 = clk::now();
  // The reference's two passes (src/main.cpp:490-515 density estimation into
  // a texture, src/main.cpp:522-535 tonemap of that texture), timed apart on
  // the device.  Without density estimation the radius-0 splat copies each
  // bin into the render-target layout.  rf_density_tonemap fuses both.
  rf_de_params none{0, 0, 0.0f};
  HIPCHK(ctx, ctx->render_img.alloc((size_t)rp->width * rp->height * 16));
  if (!ctx->ev[0]) { HIPCHK(ctx, hipEventCreate(&ctx->ev[0])); HIPCHK(ctx, hipEventCreate(&ctx->ev[1])); }
  rfk::DEArgs da;
  if ((rc = de_args(ctx, h, rp->density_estimation ? &rp->de : &none, da))) return rc;
  da.de_out = (float4*)ctx->render_img.p;
  da.do_tonemap = 0;
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  HIPCHK(ctx, rfk::launch_de_tonemap(da, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(ctx, rfk::launch_tonemap((const float4*)ctx->render_img.p, (float4*)ctx->render_img.p,
                                  rp->width * rp->height, rp->tm.gamma, rp->tm.brightness, rp->tm.vibrancy,
                                  rp->tm.scale_constant, ctx->stream));
  if (!ctx->ev[2]) { HIPCHK(ctx, hipEventCreate(&ctx->ev[2])); HIPCHK(ctx, hipEventCreate(&ctx->ev[3])); }
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(ctx->ev[2]));
  float de_ms = 0.0f, tm_ms = 0.0f;
  HIPCHK(ctx, hipEventElapsedTime(&de_ms, ctx->ev[0], ctx->ev[1]));
  HIPCHK(ctx, hipEventElapsedTime(&tm_ms, ctx->ev[1], ctx->ev[2]));
  if (hist_out) if ((rc = rf_hist_download(h, hist_out))) return rc;
  if (image_out) {
    HIPCHK(ctx, hipMemcpy(image_out, ctx->render_img.p, (size_t)rp->width * rp->height * 16, hipMemcpyDeviceToHost));
  }
  s.iterations = ctx->sp.total_particles * ((uint64_t)rp->warmup_passes + 1 + (uint64_t)rp->draw_passes * rp->draw_calls);
  s.ms_warmup = ms(t0, t1);
  s.ms_draw = ms(t1, t2);
  s.ms_density = de_ms;
  s.ms_tonemap = tm_ms;
  if (st) *st = s;
  return RF_OK;
}

int rf_particles_download(rf_ctx* ctx, rf_flame* f, float* host) {
  if (!ctx || !f || !host || !f->local.p) return set_err(ctx, RF_ERR_ARG, "no particles");
  HIPCHK(ctx, hipMemcpy(host, f->local.p, ctx->sp.total_particles * 16, hipMemcpyDeviceToHost));
  return RF_OK;
}
int rf_rng_download(rf_ctx* ctx, uint32_t* host) {
  if (!ctx || !host) return RF_ERR_ARG;
  int rc = ensure_compat(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host, ctx->rs.p, ctx->sp.total_particles * 16, hipMemcpyDeviceToHost));
  return RF_OK;
}
int rf_shuffle_download(rf_ctx* ctx, uint32_t* host) {
  if (!ctx || !host) return RF_ERR_ARG;
  int rc = ensure_compat(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host, ctx->shuf.p, ctx->shuf.bytes, hipMemcpyDeviceToHost));
  return RF_OK;
}
int rf_sample_points_download(rf_ctx* ctx, float* host) {
  if (!ctx || !host) return RF_ERR_ARG;
  int rc = ensure_compat(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host, ctx->samples.p, ctx->samples.bytes, hipMemcpyDeviceToHost));
  return RF_OK;
}

int rf_fast_perms_download(rf_ctx* ctx, uint16_t* host) {
  if (!ctx || !host) return RF_ERR_ARG;
  int rc = ensure_perms(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host, ctx->perms.p, ctx->perms.bytes, hipMemcpyDeviceToHost));
  return RF_OK;
}

int rf_eval_xform(rf_ctx* ctx, const rf_xform* xf, const float* pts, const float* rands, float* out, uint32_t n) {
  if (!ctx || !xf || !pts || !rands || !out) return set_err(ctx, RF_ERR_ARG, "null argument");
  rf_flame_desc* d = new rf_flame_desc();
  memset(d, 0, sizeof(*d));
  d->n_xforms = 1;
  d->xforms[0] = *xf;
  d->xforms[0].weight = 1.0f;
  Block b;
  std::string err;
  bool ok = build_block(*d, b, err);
  delete d;
  if (!ok) return set_err(ctx, RF_ERR_ARG, err);
  DevBuf db, dp, dr, dout;
  HIPCHK(ctx, db.alloc(sizeof(Block)));
  HIPCHK(ctx, dp.alloc((size_t)n * 12 + 16));
  HIPCHK(ctx, dr.alloc((size_t)n * 64 + 16));
  HIPCHK(ctx, dout.alloc((size_t)n * 16 + 16));
  HIPCHK(ctx, hipMemcpy(db.p, &b, sizeof(Block), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(dp.p, pts, (size_t)n * 12, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(dr.p, rands, (size_t)n * 64, hipMemcpyHostToDevice));
  HIPCHK(ctx, rfk::launch_eval_xform((const Block*)db.p, (const float*)dp.p, (const float*)dr.p, (float*)dout.p, n, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(out, dout.p, (size_t)n * 16, hipMemcpyDeviceToHost));
  return RF_OK;
}

size_t rf_struct_size(int which) {
  switch (which) {
    case 0: return sizeof(rf_var_term);
    case 1: return sizeof(rf_xform);
    case 2: return sizeof(rf_flame_desc);
    case 3: return sizeof(rf_sim_params);
    case 4: return sizeof(rf_de_params);
    case 5: return sizeof(rf_tonemap_params);
    case 6: return sizeof(rf_render_params);
    case 7: return sizeof(rf_stats);
    case 8: return sizeof(rf_frame_desc);
    case 9: return sizeof(rf_batch_config);
    default: return 0;
  }
}

int rf_atomic_probe(rf_ctx* ctx, int32_t hist_mode, uint64_t footprint_bins, uint64_t samples, double* ms_out) {
  if (!ctx || !ms_out || footprint_bins == 0) return set_err(ctx, RF_ERR_ARG, "bad probe arguments");
  if (footprint_bins > (1ull << 28)) return set_err(ctx, RF_ERR_LIMIT, "probe footprint above 2^28 bins");
  if (samples < 64 * 256 || samples % (64 * 256) != 0) return set_err(ctx, RF_ERR_ARG, "samples must be a multiple of 16384");
  HIPCHK(ctx, ctx->probe.alloc(footprint_bins * 16 + 256 * 16));
  float4* pal = (float4*)((char*)ctx->probe.p + footprint_bins * 16);
  HIPCHK(ctx, hipMemsetAsync(ctx->probe.p, 0, footprint_bins * 16 + 256 * 16, ctx->stream));
  if (!ctx->ev[0]) { HIPCHK(ctx, hipEventCreate(&ctx->ev[0])); HIPCHK(ctx, hipEventCreate(&ctx->ev[1])); }
  HIPCHK(ctx, rfk::launch_atomic_probe((float*)ctx->probe.p, footprint_bins, samples, hist_mode, pal, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  HIPCHK(ctx, rfk::launch_atomic_probe((float*)ctx->probe.p, footprint_bins, samples, hist_mode, pal, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  *ms_out = ms;
  return RF_OK;
}

}  // extern "C"
