// Device-side flame representation, random number generators and the xform
// interpreter shared by the compat and fast chaos-game engines.
//
// The reference compiles one GLSL `dispatch(v, xform)` per flame
// (src/variation_table.cpp:217-265).  Here a flame is data: a `Block` in
// device memory, read with wave-uniform addresses (scalar loads), and each
// wave runs one xform at a time so the variation `switch` is uniform.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace rfd {

constexpr int MAX_XF = 32;          // non-final xforms per flame
constexpr int FINAL = MAX_XF;       // slot of the final xform in Block::xf
constexpr int MAX_TERMS = 256;      // variation terms per flame (all xforms)
constexpr int MAX_PRM = 8;

constexpr float PI = 3.141592653589793f;   // shaders/include/math.glsl:1
constexpr float PI_2 = PI / 2.0f;
constexpr float EPS = 1e-10f;              // shaders/include/math.glsl:4
constexpr float M_2_PI_F = 0.636619772367581343f;
constexpr float LOG10E = 0.434294481903251827651128918916f;  // math.glsl log10

// precalc bits (must match gen_variations.py)
constexpr uint32_t PC_R = 1u << 0, PC_RSQ = 1u << 1, PC_A = 1u << 2, PC_PHI = 1u << 3,
                   PC_SINA = 1u << 4, PC_COSA = 1u << 5, PC_SINR = 1u << 6, PC_COSR = 1u << 7;
// term flags (variations.yaml flags)
constexpr uint32_t TF_NO_WEIGHT_MUL = 1u, TF_PRE_XFORM = 2u;
// xform flags
constexpr uint32_t XF_HAS_POST = 1u;

struct Term {            // 48 bytes
  uint32_t var;          // variation id
  uint32_t flags;        // TF_*
  float weight;
  float prm[MAX_PRM];
  uint32_t pad;
};

struct Xform {           // 88 bytes
  float aff[6];          // a b c d e f ; index = col*2 + row (c00 c01 c10 c11 c20 c21)
  float post[6];
  float color, color_speed, opacity;
  // the colour blend's uniform factors, rounded as rf_mix rounds them:
  // c_keep = 1 - color_speed, c_add = color * color_speed
  float c_keep, c_add;
  uint32_t flags;
  uint32_t precalc;      // union of the regular terms' precalc needs
  uint32_t term_begin;   // first term in Block::terms
  uint32_t n_pre;        // leading pre_xform terms
  uint32_t n_terms;      // all terms
};

struct Block {
  uint32_t n_xforms;     // non-final xforms (>= 1)
  uint32_t has_final;
  uint32_t n_terms;
  uint32_t pad;
  float cdf[MAX_XF];     // running sum of normalised weights, in xform order
  // For a raw 32-bit draw x with u = float(x) * 2^-32:  cdf[i] >= u  <=>
  // (never_mask bit i clear and x <= cdf_lim[i]).  Lets a wave select on SALU.
  uint32_t cdf_lim[MAX_XF];
  uint32_t never_mask;
  uint32_t pad2[3];
  Xform xf[MAX_XF + 1];
  Term terms[MAX_TERMS];
};

// A Block seen through the constant address space: it is never written while
// a chaos-game kernel runs, and loads through such a pointer are always
// scalar (s_load) when their address is uniform, so the flame is read through
// the scalar cache instead of vector loads that would queue behind atomics.
typedef const __attribute__((address_space(4))) Block CBlock;
__device__ __forceinline__ CBlock* as_constant(const Block* p) { return (CBlock*)p; }

// ---------------------------------------------------------------- helpers
__host__ __device__ __forceinline__ float rf_mod2(float x, float y) {
  return x - y * truncf(x / y);  // shaders/include/math.glsl mod2
}
__host__ __device__ __forceinline__ bool rf_badval(float x) {
  return (x != x) || (x > 1e10f) || (x < -1e10f);  // shaders/include/math.glsl badval
}

// sin/cos on [0, 1] by their Taylor series in Horner form (truncation error
// < 2e-10), written with explicit fmaf so the CPU oracle and the GPU produce
// identical bits.  Used for the first-run jitter angle, which is sqrt(U[0,1]).
__host__ __device__ __forceinline__ void rf_sincos01(float q, float* s, float* c) {
  float z = q * q;
  float ps = -2.5052108385441720e-08f;
  ps = fmaf(ps, z, 2.7557319223985893e-06f);
  ps = fmaf(ps, z, -1.9841269841269841e-04f);
  ps = fmaf(ps, z, 8.3333333333333332e-03f);
  ps = fmaf(ps, z, -1.6666666666666666e-01f);
  ps = fmaf(ps, z, 1.0f);
  *s = q * ps;
  float pc = 2.0876756987868099e-09f;
  pc = fmaf(pc, z, -2.7557319223985888e-07f);
  pc = fmaf(pc, z, 2.4801587301587302e-05f);
  pc = fmaf(pc, z, -1.3888888888888889e-03f);
  pc = fmaf(pc, z, 4.1666666666666664e-02f);
  pc = fmaf(pc, z, -0.5f);
  *c = fmaf(pc, z, 1.0f);
}

// ------------------------------------------------------------------ math
// Transcendentals of the variation interpreter.  The device library's sinf
// and friends inline a Payne-Hanek reduction at every call site, and the
// compiler if-converts it, so every call paid for both paths; in a kernel with
// ~100 trig call sites that multiplied the iteration's VALU work ~8x.
//
// sin/cos: the argument in revolutions, x / (2 pi) rounded to float, reduced
// to [0, 1) with v_fract_f32, then the hardware v_sin_f32 / v_cos_f32 (whose
// own input range ends at 256 revolutions).  The rounding of x / (2 pi) costs
// at most 2^-25 |x| radians, so the error stays within a few 1e-7 on
// [-pi, pi] (GLSL asks 2^-11 there and nothing outside it) and grows as
// 6e-8 |x| beyond.  Four instructions for both values, against eight with
// a three-part Cody-Waite reduction and about 25 for a polynomial pair.
__device__ __forceinline__ void rf_sincos(float x, float* s, float* c) {
  const float t = __builtin_amdgcn_fractf(x * 0.159154936671257019f);
  *s = __builtin_amdgcn_sinf(t);
  *c = __builtin_amdgcn_cosf(t);
}
__device__ __forceinline__ float rf_sin(float x) { float s, c; rf_sincos(x, &s, &c); return s; }
__device__ __forceinline__ float rf_cos(float x) { float s, c; rf_sincos(x, &s, &c); return c; }
__device__ __forceinline__ float rf_tan(float x) { float s, c; rf_sincos(x, &s, &c); return s * __builtin_amdgcn_rcpf(c); }
// exp, log and pow through the hardware base-2 instructions; GLSL defines
// pow(x, y) as exp2(y * log2(x)) for x > 0.
__device__ __forceinline__ float rf_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float rf_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float rf_exp(float x) { return rf_exp2(x * 1.44269504088896341f); }
__device__ __forceinline__ float rf_log(float x) { return rf_log2(x) * 0.693147180559945309f; }
__device__ __forceinline__ float rf_log10(float x) { return rf_log2(x) * 0.301029995663981195f; }
__device__ __forceinline__ float rf_pow(float x, float y) { return rf_exp2(y * rf_log2(x)); }
__device__ __forceinline__ float rf_sinh(float x) {
  const float e = rf_exp(x);
  return 0.5f * (e - __builtin_amdgcn_rcpf(e));
}
__device__ __forceinline__ float rf_cosh(float x) {
  const float e = rf_exp(x);
  return 0.5f * (e + __builtin_amdgcn_rcpf(e));
}
// Division and square root of the variation interpreter: the hardware
// reciprocal (1 ulp) times the dividend, and the hardware square root, both
// within the 2.5 ulp GLSL allows for division and sqrt.  The compiler's own
// 2.5 ulp division rescales both operands with frexp/ldexp to cover denormal
// and huge divisors (8 instructions per division instead of 2); chaos-game
// coordinates stay within +-1e10 (larger ones are reset), so the rescaling
// buys nothing here.
__device__ __forceinline__ float rf_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rf_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// atan2(y, x): the quotient t = min(|x|,|y|) / max(|x|,|y|) in [0, 1], an odd
// polynomial for atan(t) (relative-error minimax fit of atan(t)/t in t^2 on
// [0, 1], 8 terms, max error 2.9 ulp in float32 evaluation), then the octant
// reflections pi/2 - r and pi - r and the sign of y.  The sign bit of x picks
// the pi - r branch, so atan2(+-0, -0) = +-pi as in C.  About 20 instructions
// against about 40 for the device library's atan2f, whose infinity handling
// the interpreter does not need (coordinates are finite).
__device__ __forceinline__ float rf_atan2(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.0f ? mn * rf_rcp(mx) : 0.0f;
  const float z = t * t;
  float p = -4.693436035e-03f;
  p = fmaf(p, z, 2.425300952e-02f);
  p = fmaf(p, z, -5.948731153e-02f);
  p = fmaf(p, z, 9.914363580e-02f);
  p = fmaf(p, z, -1.401951006e-01f);
  p = fmaf(p, z, 1.996973005e-01f);
  p = fmaf(p, z, -3.333199132e-01f);
  p = fmaf(p, z, 9.999999011e-01f);
  float r = t * p;
  if (ay > ax) r = PI_2 - r;
  if (__float_as_uint(x) >> 31) r = PI - r;
  return __uint_as_float((__float_as_uint(r) & 0x7fffffffu) | (__float_as_uint(y) & 0x80000000u));
}

// Correctly rounded sqrt and division for the few places whose result must
// match the CPU bit for bit (the library is built with fast fp32 div/sqrt).
// They go through binary64: rounding the correctly rounded double result to
// float is exact-rounding for sqrt and division because 53 >= 2*24 + 2
// (Figueroa, "When is double rounding innocuous?", SIGNUM 1995).
__device__ __forceinline__ float rf_sqrt_rn(float x) { return (float)sqrt((double)x); }
__device__ __forceinline__ float rf_div_rn(float x, float y) { return (float)((double)x / (double)y); }

// Palette index min(255, uint(ceil(c * 255))) (shaders/flame.glsl:84), with
// negative and NaN colours mapped to 0 (undefined in GLSL).
__host__ __device__ __forceinline__ uint32_t rf_pal_index(float c) {
  // ceil(c * 255) clamped to [0, 255], NaN to 0, without branches: fmaxf
  // returns its non-NaN operand
  const float t = fminf(fmaxf(ceilf(c * 255.0f), 0.0f), 255.0f);
  return (uint32_t)t;
}

// GLSL mix(x, y, a) = x*(1-a) + y*a, evaluated without contraction.
__host__ __device__ __forceinline__ float rf_mix(float x, float y, float a) {
  float t0 = x * (1.0f - a);
  float t1 = y * a;
  return t0 + t1;
}

// ------------------------------------------------------------------ RNGs
// Bob Jenkins' small fast PRNG (JSF32), the generator of the reference
// (src/util.hpp:75-96, shaders/include/random.glsl:29-41).  The GPU side of
// the reference returns the new `a`; the CPU seeding returns `d`.
struct Jsf32 {
  uint32_t a, b, c, d;
  __host__ __device__ __forceinline__ static uint32_t rot(uint32_t x, int k) {
    return (x << k) | (x >> (32 - k));
  }
  __host__ __device__ __forceinline__ void step() {
    uint32_t e = a - rot(b, 27);
    a = b ^ rot(c, 17);
    b = c + d;
    c = d + e;
    d = e + a;
  }
  __host__ __device__ __forceinline__ uint32_t next_a() { step(); return a; }  // random.glsl:30-37
  __host__ __device__ __forceinline__ uint32_t next_d() { step(); return d; }  // util.hpp:81-88
  // randf(): clamp(float(u) / 4294967295.0f, 0, 1); the literal rounds to 2^32
  // so this is float(u) * 2^-32 exactly (random.glsl:39-41).
  __host__ __device__ __forceinline__ float randf() { return (float)next_a() * 0x1p-32f; }
  // jsf32::warmup_ctx (src/util.hpp:90-95)
  __host__ __device__ static Jsf32 seeded(uint32_t seed) {
    Jsf32 s{0xf1ea5eedu, seed, seed, seed};
    for (int i = 0; i < 20; ++i) s.step();
    return s;
  }
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Counter-based: the fast engine derives every
// lane's and wave's stream from (seed, particle/wave id, launch) with it.
struct Philox {
  __host__ __device__ static inline void round(uint32_t ctr[4], const uint32_t key[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ ctr[1] ^ key[0];
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ ctr[3] ^ key[1];
    uint32_t n3 = lo0;
    ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
  }
  __host__ __device__ static inline void gen(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c[4] = {in[0], in[1], in[2], in[3]};
    uint32_t k[2] = {key_in[0], key_in[1]};
    for (int i = 0; i < 10; ++i) {
      if (i) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
      round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
  }
};

// A JSF32 state seeded from a Philox block, warmed like warmup_ctx.
__host__ __device__ inline Jsf32 jsf_from_philox(uint64_t seed, uint32_t id, uint32_t launch, uint32_t stream) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t ctr[4] = {id, launch, stream, 0x5EEDu};
  uint32_t o[4];
  Philox::gen(ctr, key, o);
  Jsf32 s{0xf1ea5eedu ^ o[0], o[1], o[2], o[3] | 1u};
  for (int i = 0; i < 8; ++i) s.step();
  return s;
}

// Random source backed by a table (diagnostic kernel and tests).
struct TableRng {
  const float* p;
  int k;
  __device__ __forceinline__ float randf() { return p[k++]; }
};

// ----------------------------------------------------------- interpreter
struct Pre {
  float r, rsq, a, phi, sina, cosa, sinr, cosr;
};

template <class Rng>
__device__ __forceinline__ void eval_var(uint32_t id, float x, float y, const Pre& pc, float w,
                                         const float* __restrict__ A, const float* __restrict__ P,
                                         Rng& rng, float& ox, float& oy) {
  const float r = pc.r, rsq = pc.rsq, a = pc.a, phi = pc.phi;
  const float sina = pc.sina, cosa = pc.cosa, sinr = pc.sinr, cosr = pc.cosr;
  (void)r; (void)rsq; (void)a; (void)phi; (void)sina; (void)cosa; (void)sinr; (void)cosr;
  (void)w; (void)A; (void)P;
  switch (id) {
#include "gen/variation_cases.inc"
    default: ox = 0.0f; oy = 0.0f; break;
  }
}

// Result of one xform application: the new point, colour and opacity
// (dispatch() in the generated GLSL returns vec4(result, color, opacity)).
struct XOut { float x, y, c, w; };

// The pieces of one xform application, shared by the interpreter
// (apply_xform: structure read from the Block at run time) and by compiled
// flame programs (apply_fixed: structure fixed at compile time, coefficients
// still read from the Block), so both evaluate the same operations in the
// same order and produce the same bits.
// Evaluation order follows variation_table.cpp:97-165: fma affine, pre_xform
// terms added into v, precalcs, the regular terms summed in alphabetical order
// (separate multiply and add), then the fma post-affine.

// One pre_xform term: its (weighted) result is added into the point.
template <class Rng, class TP>
__device__ __forceinline__ void pre_term(uint32_t var, uint32_t flags, TP t, float& x, float& y, const Pre& pc,
                                         const float* A, Rng& rng) {
  float P[MAX_PRM];
#pragma unroll
  for (int j = 0; j < MAX_PRM; ++j) P[j] = t->prm[j];
  const float w = t->weight;
  float ox, oy;
  eval_var(var, x, y, pc, w, A, P, rng, ox, oy);
  if (!(flags & TF_NO_WEIGHT_MUL)) { ox = w * ox; oy = w * oy; }
  x = x + ox;
  y = y + oy;
}

// The precalcs named by mask m (PC_* bits); the others stay 0.
__device__ __forceinline__ void precalc(uint32_t m, float x, float y, Pre& pc) {
  if (m & (PC_RSQ | PC_R)) pc.rsq = x * x + y * y;
  if (m & PC_R) pc.r = rf_sqrt(pc.rsq);
  if (m & PC_A) pc.a = rf_atan2(x, y);
  if (m & PC_PHI) pc.phi = rf_atan2(y, x);
  if (m & (PC_SINA | PC_COSA)) {
    const float ir = rf_rcp(pc.r);
    pc.sina = x * ir;
    pc.cosa = y * ir;
  }
  if (m & (PC_SINR | PC_COSR)) rf_sincos(pc.r, &pc.sinr, &pc.cosr);
}

// One regular term: the first one sets the sum, the others add to it.
template <class Rng, class TP>
__device__ __forceinline__ void reg_term(uint32_t var, uint32_t flags, TP t, float x, float y, const Pre& pc,
                                         const float* A, Rng& rng, bool first_term, float& rx, float& ry) {
  float P[MAX_PRM];
#pragma unroll
  for (int j = 0; j < MAX_PRM; ++j) P[j] = t->prm[j];
  const float w = t->weight;
  float ox, oy;
  eval_var(var, x, y, pc, w, A, P, rng, ox, oy);
  if (!(flags & TF_NO_WEIGHT_MUL)) { ox = w * ox; oy = w * oy; }
  if (first_term) { rx = ox; ry = oy; }
  else { rx = rx + ox; ry = ry + oy; }
}

// Post-affine (when `post`) and colour blend.  If `first` the incoming
// colour is replaced by a random draw made after the variations' draws, as
// the generated code does on the first run (variation_table.cpp:241-243).
template <class Rng, class XP>
__device__ __forceinline__ XOut finish_xform(XP xf, bool post, float rx, float ry, float c0, bool first, Rng& rng) {
  if (post) {
    const float q0 = xf->post[0], q1 = xf->post[1], q2 = xf->post[2];
    const float q3 = xf->post[3], q4 = xf->post[4], q5 = xf->post[5];
    float px = fmaf(q0, rx, fmaf(q2, ry, q4));
    float py = fmaf(q1, rx, fmaf(q3, ry, q5));
    rx = px; ry = py;
  }
  float cin = first ? rng.randf() : c0;
  XOut o;
  o.x = rx; o.y = ry;
  // rf_mix(cin, color, color_speed) with its uniform factors precomputed
  const float t0 = cin * xf->c_keep;
  o.c = t0 + xf->c_add;
  o.w = xf->opacity;
  return o;
}

// Applies xform `xi` of `blk` to (x0, y0, c0), structure read from the Block.
template <class Rng, class BP>
__device__ __forceinline__ XOut apply_xform(BP blk, int xi, float x0, float y0,
                                            float c0, bool first, Rng& rng) {
  const auto* __restrict__ xf = &blk->xf[xi];
  float A[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) A[k] = xf->aff[k];
  float x = fmaf(A[0], x0, fmaf(A[2], y0, A[4]));
  float y = fmaf(A[1], x0, fmaf(A[3], y0, A[5]));
  const uint32_t tb = xf->term_begin, npre = xf->n_pre, nt = xf->n_terms;
  const auto* __restrict__ T = blk->terms + tb;
  Pre pc;
  pc.r = pc.rsq = pc.a = pc.phi = pc.sina = pc.cosa = pc.sinr = pc.cosr = 0.0f;
  for (uint32_t k = 0; k < npre; ++k) pre_term(T[k].var, T[k].flags, T + k, x, y, pc, A, rng);
  precalc(xf->precalc, x, y, pc);
  float rx = 0.0f, ry = 0.0f;
  for (uint32_t k = npre; k < nt; ++k) {
    reg_term(T[k].var, T[k].flags, T + k, x, y, pc, A, rng, k == npre, rx, ry);
    // Everything a variation computes from x, y, the precalcs or the affine is
    // invariant in this loop, and loop-invariant code motion would hoist it out
    // of the switch -- i.e. evaluate sin(x), sin(rsq), exp(x - 1), ... of all 76
    // variations on every xform application.  An empty asm that claims to
    // rewrite these values at the end of every term makes them loop-carried,
    // which keeps each case's work inside its case; it rewrites the values in
    // place, so it costs no register copies.
    asm volatile("" : "+v"(x), "+v"(y), "+v"(pc.r), "+v"(pc.rsq), "+v"(pc.a), "+v"(pc.phi), "+v"(pc.sina),
                 "+v"(pc.cosa), "+v"(pc.sinr), "+v"(pc.cosr));
    asm volatile("" : "+s"(A[0]), "+s"(A[1]), "+s"(A[2]), "+s"(A[3]), "+s"(A[4]), "+s"(A[5]));
  }
  return finish_xform(xf, (xf->flags & XF_HAS_POST) != 0u, rx, ry, c0, first, rng);
}

// A list of term codes (variation id | TF_* flags << 16) fixed at compile time.
template <uint32_t... C>
struct TermList {};

// Applies xform `xi` whose structure is fixed at compile time: the precalc
// mask PCM, whether it has a post-affine, the index of its first term in the
// Block (TB; TERM_BEGIN_LOAD: read it from the Block) and its pre_xform and
// regular terms in Block order.  Coefficients, weights and parameters are
// still read from the Block, so a flame's values can change without a new
// program; with TB fixed, none of those loads waits for another.
constexpr uint32_t TERM_BEGIN_LOAD = 0xffffffffu;
template <uint32_t PCM, bool POST, uint32_t TB, class Rng, class BP, uint32_t... PRE, uint32_t... REG>
__device__ __forceinline__ XOut apply_fixed(TermList<PRE...>, TermList<REG...>, BP blk, int xi, float x0,
                                            float y0, float c0, bool first, Rng& rng) {
  const auto* __restrict__ xf = &blk->xf[xi];
  float A[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) A[k] = xf->aff[k];
  float x = fmaf(A[0], x0, fmaf(A[2], y0, A[4]));
  float y = fmaf(A[1], x0, fmaf(A[3], y0, A[5]));
  const auto* __restrict__ T = blk->terms + (TB == TERM_BEGIN_LOAD ? xf->term_begin : TB);
  Pre pc;
  pc.r = pc.rsq = pc.a = pc.phi = pc.sina = pc.cosa = pc.sinr = pc.cosr = 0.0f;
  uint32_t k = 0;
  ((pre_term(PRE & 0xffffu, PRE >> 16, T + k, x, y, pc, A, rng), ++k), ...);
  precalc(PCM, x, y, pc);
  float rx = 0.0f, ry = 0.0f;
  bool ft = true;
  ((reg_term(REG & 0xffffu, REG >> 16, T + k, x, y, pc, A, rng, ft, rx, ry), ++k, ft = false), ...);
  return finish_xform(xf, POST, rx, ry, c0, first, rng);
}

// get_xform_id (shaders/templates/xform_select.tpl.glsl:1-14): the first i
// whose running weight sum reaches u; the last xform is unconditional.
template <class BP>
__device__ __forceinline__ int select_xform(BP blk, float u) {
  const int n = (int)blk->n_xforms;
  int i = 0;
  while (i < n - 1 && !(blk->cdf[i] >= u)) ++i;
  return i;
}

// The same selection for a wave-uniform raw draw x (u = float(x) * 2^-32),
// entirely on the scalar unit: gfx950 has no scalar float compare, so the
// host precomputes integer limits with cdf[i] >= u <=> x <= cdf_lim[i].
template <class BP>
__device__ __forceinline__ int select_xform_raw(BP blk, uint32_t x_any) {
  const uint32_t x = __builtin_amdgcn_readfirstlane(x_any);
  const int n = (int)blk->n_xforms;
  const uint32_t never = blk->never_mask;
  int i = 0;
  while (i < n - 1 && !(((never >> i) & 1u) == 0u && x <= blk->cdf_lim[i])) ++i;
  return i;
}

// select_xform_raw for flames of at most NMAX xforms: the first NMAX - 1
// limits are compared independently (their scalar loads issue together, no
// load-compare-branch chain) into a mask of the xforms that would accept x,
// and the lowest accepting one is taken; the same index as the loop.
template <int NMAX, class BP>
__device__ __forceinline__ int select_xform_upto(BP blk, uint32_t x_any) {
  static_assert(NMAX >= 1 && NMAX <= MAX_XF, "xform count");
  const uint32_t x = __builtin_amdgcn_readfirstlane(x_any);
  const uint32_t n = blk->n_xforms;  // 1 <= n <= NMAX
  uint32_t ok = 0u;
#pragma unroll
  for (int i = 0; i < NMAX - 1; ++i) ok |= (x <= blk->cdf_lim[i] ? 1u : 0u) << i;
  const uint32_t last = 1u << (n - 1u);  // the last xform is unconditional
  ok = (ok & ~blk->never_mask & (last - 1u)) | last;
  return (int)__builtin_ctz(ok);
}

}  // namespace rfd
