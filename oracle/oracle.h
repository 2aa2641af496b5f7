/*
 * CPU oracle for refrakt_amd — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference renderer's hot path
 * (untrioctium/refrakt: shaders/flame.glsl, shaders/include/random.glsl,
 * shaders/templates/{xform_select,animate}.tpl.glsl, variations.yaml,
 * src/flame.cpp, src/util.hpp, src/hammersley.cpp,
 * shaders/density_{vert,frag}.glsl, shaders/tonemap.glsl, src/main.cpp).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker; the library never calls it.
 *
 * Parts and how each is pinned (tests/test_oracle.py):
 *   jsf32 .............. cross-implementation from Jenkins' published definition,
 *                         the GLSL/CPU output relation d = e + a, invertibility
 *   philox4x32-10 ....... Random123 known-answer vectors
 *   hammersley .......... brute-force radical inverse
 *   xform selection ..... boundary cases and chi-square frequencies
 *   variations (f64) .... closed-form identities and hand-computed points
 *   compat pass (f32) ... closed-form trajectories of affine flames, invariants
 *                         (permutation conservation, hit counts = drawn)
 *   fast replica (f32) .. same invariants; it restates this project's own
 *                         schedule (wave-uniform xforms + LDS permutations), not
 *                         a reference algorithm
 *   chaos f64 ........... fixed points / closed-form attractors (single-map and
 *                         two-point IFS), mass conservation
 *   density estimation .. identity at r=0, closed-form stamp sums, symmetry,
 *                         scatter == brute-force gather
 *   tonemap ............. closed-form pixels (vibrancy 0 and 1)
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_XF 32
#define ORC_MAX_TERMS 16
#define ORC_MAX_PRM 8

/* ---------------------------------------------------------------- RNGs */
typedef struct orc_jsf { uint32_t a, b, c, d; } orc_jsf;
void orc_jsf_seed(orc_jsf* s, uint32_t seed);          /* util.hpp:90-95 */
uint32_t orc_jsf_next_a(orc_jsf* s);                     /* random.glsl:30-37 */
uint32_t orc_jsf_next_d(orc_jsf* s);                     /* util.hpp:81-88 */
float orc_jsf_randf(orc_jsf* s);                         /* random.glsl:39-41 */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_jsf_from_philox(orc_jsf* s, uint64_t seed, uint32_t id, uint32_t launch, uint32_t stream);

/* Random source for the variation functions. */
typedef struct orc_rng {
  int kind;              /* 0: jsf32 GLSL variant, 1: table */
  orc_jsf* jsf;
  const float* table;
  int pos;
} orc_rng;
double orc_rng_next(orc_rng* r);

/* ------------------------------------------------------------ model */
int orc_var_count(void);
const char* orc_var_name(int id);
int orc_var_id(const char* name);
int orc_var_nparams(int id);
const char* orc_var_param(int id, int k);

typedef struct orc_term { int32_t var; float weight; float prm[ORC_MAX_PRM]; } orc_term;
typedef struct orc_xform {
  float aff[6]; float post[6]; int32_t has_post;
  int32_t n_terms; orc_term terms[ORC_MAX_TERMS];   /* any order: sorted by name internally */
  float weight, color, color_speed, opacity, rotation_frequency;
} orc_xform;
typedef struct orc_flame {
  int32_t n_xforms; orc_xform xf[ORC_MAX_XF];
  int32_t has_final; orc_xform fin;
  float pal[256][4];
} orc_flame;

/* Variation k of the reference at (x, y) in double precision; w is the
 * term weight, aff the xform's affine (for the $cXY references). */
void orc_var_eval(int id, double x, double y, double w, const float* prm, const float* aff, orc_rng* rng,
                  double* ox, double* oy);
/* One xform (affine, pre_xform terms, precalcs, terms in name order, post,
 * colour blend).  f32: the reference's float pipeline (fma affine, float
 * sums; each variation value computed in double and rounded).  f64: all double. */
void orc_apply_f32(const orc_xform* xf, float x, float y, float c, int first, orc_rng* rng, float out[4]);
void orc_apply_f64(const orc_xform* xf, double x, double y, double c, int first, orc_rng* rng, double out[4]);
/* get_xform_id over normalised float weights (xform_select.tpl.glsl). */
int orc_select(const orc_flame* f, float u);
void orc_cdf(const orc_flame* f, float cdf[ORC_MAX_XF]);
/* animate.tpl.glsl for temporal sample i of T. */
void orc_animate(const orc_flame* src, uint32_t i, uint32_t T, float tss, orc_flame* dst);
/* Screen affine of draw_to_bins (flame.cpp:289-296). */
void orc_screen_affine(uint32_t W, uint32_t H, const uint32_t size[2], const float center[2], float scale,
                       float rotate, float out[6]);
/* Hammersley points (hammersley.cpp:29-48): count x 4 floats. */
void orc_hammersley(uint32_t count, float* out);
int orc_pal_index(float c);

/* ---------------------------------------------------- chaos game (f32) */
/* One pass of the reference iterate shader (flame.glsl:41-90) over
 * P = ppt * T particles.  flames: one flame (stride 0) or T flames.
 * bins (W*H*4 doubles) and hits (W*H) may be NULL. */
void orc_compat_pass(const orc_flame* flames, int stride, uint32_t T, uint32_t ppt,
                     const float* pos_in, float* pos_out, const uint32_t* shuf, uint32_t* rs,
                     uint32_t in_row, uint32_t out_row, int random_read, int random_write, int first_run,
                     int do_draw, uint32_t W, uint32_t H, const float ss[6],
                     double* bins, uint32_t* hits, uint64_t* drawn);

/* Seeded jsf32 states of slots 0..n-1 (orc_jsf_seed each), n x 4 words. */
void orc_jsf_seed_all(uint32_t* out, uint64_t n);
/* n_passes reference passes over P = ppt * T particles on `threads` threads
 * (workgroups split between threads, a barrier after each pass), ping-pong
 * between pos_a and pos_b (pass k reads pos[k % 2]), shuffle rows
 * rows[2k], rows[2k + 1]; bins: W*H*4 floats added without atomics like the
 * shader (may be NULL when do_draw is 0).  Returns the binned count. */
uint64_t orc_ref_passes_mt(const orc_flame* f, uint32_t T, uint32_t ppt, float* pos_a, float* pos_b,
                           const uint32_t* shuf, uint32_t* rs, const uint32_t* rows, uint32_t n_passes,
                           int random_read, int random_write, int first_run, int do_draw, uint32_t W, uint32_t H,
                           const float ss[6], float* bins, int threads);

/* This project's fast engine schedule (not a reference algorithm): workgroups
 * of 64*waves particles, one xform per wave per iteration, LDS permutation
 * rows from `perms` (ORC_NPERM rows x 64*waves, row = stream & (ORC_NPERM - 1)). */
#define ORC_NPERM 8u
void orc_fast_run(const orc_flame* flames, int stride, uint32_t ppt, uint64_t P, int waves,
                  const uint16_t* perms, uint64_t seed, uint32_t launch, uint32_t n_iter, int first_run,
                  int do_draw, uint32_t W, uint32_t H, const float ss[6], float* pos,
                  double* bins, uint32_t* hits, uint64_t* drawn);

/* ----------------------------------------------- chaos game (f64 oracle) */
/* Independent per-particle chaos game in double precision: every particle
 * draws its own xform each iteration (the textbook algorithm of Draves &
 * Reckase), starts at a uniform point of [-1,1]^2 with a uniform colour,
 * iterates `fuse` times unbinned, then `iters` times binned.  Multithreaded;
 * bins: W*H*4 doubles (zeroed by the caller).  Returns binned samples. */
uint64_t orc_chaos_f64(const orc_flame* f, uint32_t W, uint32_t H, const double ss[6],
                       uint64_t particles, uint32_t fuse, uint32_t iters, uint64_t seed, int threads,
                       double* bins);

/* ------------------------------------------------------------ DE, tonemap */
/* The reference's density-estimation splat (density_vert.glsl,
 * density_frag.glsl, main.cpp:490-515) as a scatter: hist in bin layout
 * (W*H*4 floats), out in render-target layout (W*H*4 doubles, zeroed by the
 * caller).  fp32 != 0 computes radius and weights in float. */
void orc_density(const float* hist, uint32_t W, uint32_t H, int R, int rmin, float curve, int fp32, double* out);
/* tonemap.glsl:18-39 on n RGBA pixels (double in/out). */
void orc_tonemap(const double* in, double* out, uint64_t n, double gamma, double brightness, double vibrancy,
                 double k);

#ifdef __cplusplus
}
#endif
#endif
