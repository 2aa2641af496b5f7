/*
 * CPU oracle for refrakt_amd — TEST INFRASTRUCTURE ONLY (see oracle.h for
 * what each part restates and how it is pinned).  Compiled with
 * -ffp-contract=off so float expressions round exactly as written.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static const double PI_D = 3.14159265358979323846;
static const double EPS_D = (double)1e-10f;    /* math.glsl:4, a float constant */
static const float PI_F = 3.141592653589793f;

/* ================================================================= RNGs */
static uint32_t rot32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
static void jsf_step(orc_jsf* s) {           /* util.hpp:81-88, random.glsl:30-37 */
  uint32_t e = s->a - rot32(s->b, 27);
  s->a = s->b ^ rot32(s->c, 17);
  s->b = s->c + s->d;
  s->c = s->d + e;
  s->d = e + s->a;
}
void orc_jsf_seed(orc_jsf* s, uint32_t seed) {  /* util.hpp:90-95 */
  s->a = 0xf1ea5eedu;
  s->b = s->c = s->d = seed;
  for (int i = 0; i < 20; ++i) jsf_step(s);
}
uint32_t orc_jsf_next_a(orc_jsf* s) { jsf_step(s); return s->a; }
uint32_t orc_jsf_next_d(orc_jsf* s) { jsf_step(s); return s->d; }
float orc_jsf_randf(orc_jsf* s) {              /* random.glsl:39-41 */
  float u = (float)orc_jsf_next_a(s);
  float r = u / 4294967295.0f;
  if (r < 0.0f) r = 0.0f;
  if (r > 1.0f) r = 1.0f;
  return r;
}

/* Philox4x32-10 (Salmon et al., SC'11): 10 rounds of
 * (hi(M0*c0) ^ c1 ^ k0, lo(M1*c2), hi(M1*c2) ^ c3 ^ k1, lo(M0*c0)) with
 * M0 = 0xD2511F53, M1 = 0xCD9E8D57 and Weyl key bumps 0x9E3779B9, 0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The fast engine's stream derivation: a Philox block keyed by the seed with
 * counter (id, launch, stream, 0x5EED) becomes a JSF32 state, stepped 8 times. */
void orc_jsf_from_philox(orc_jsf* s, uint64_t seed, uint32_t id, uint32_t launch, uint32_t stream) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t ctr[4] = {id, launch, stream, 0x5EEDu};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  s->a = 0xf1ea5eedu ^ o[0];
  s->b = o[1];
  s->c = o[2];
  s->d = o[3] | 1u;
  for (int i = 0; i < 8; ++i) jsf_step(s);
}

/* xoshiro256** (Blackman & Vigna) for the f64 oracle's own streams. */
static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xo_next(uint64_t s[4]) {
  uint64_t r = rotl64(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl64(s[3], 45);
  return r;
}
static uint64_t sm64(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double xo_unit(uint64_t s[4]) { return (double)(xo_next(s) >> 11) * 0x1p-53; }

typedef struct { uint64_t s[4]; } xo_state;

double orc_rng_next(orc_rng* r) {
  if (r->kind == 0) return (double)orc_jsf_randf(r->jsf);
  if (r->kind == 1) return (double)r->table[r->pos++];
  return xo_unit(((xo_state*)r->jsf)->s);
}

/* ============================================================ variations */
/* Names, parameters and flags of the reference registry (variations.yaml). */
typedef struct { const char* name; int nprm; const char* prm[ORC_MAX_PRM]; int no_weight_mul, pre_xform; } var_def;
static const var_def VARS[] = {
  {"linear", 0, {0}, 0, 0}, {"sinusoidal", 0, {0}, 0, 0}, {"spherical", 0, {0}, 0, 0}, {"swirl", 0, {0}, 0, 0},
  {"horseshoe", 0, {0}, 0, 0}, {"polar", 0, {0}, 0, 0}, {"handkerchief", 0, {0}, 0, 0}, {"heart", 0, {0}, 0, 0},
  {"disc", 0, {0}, 0, 0}, {"spiral", 0, {0}, 0, 0}, {"hyperbolic", 0, {0}, 0, 0}, {"diamond", 0, {0}, 0, 0},
  {"ex", 0, {0}, 0, 0}, {"julia", 0, {0}, 0, 0}, {"bent", 0, {0}, 0, 0}, {"waves", 0, {0}, 0, 0},
  {"fisheye", 0, {0}, 0, 0}, {"popcorn", 0, {0}, 0, 0}, {"exponential", 0, {0}, 0, 0}, {"power", 0, {0}, 0, 0},
  {"cosine", 0, {0}, 0, 0}, {"rings", 0, {0}, 0, 0}, {"fan", 0, {0}, 0, 0},
  {"blob", 3, {"blob_high", "blob_low", "blob_waves"}, 0, 0},
  {"pdj", 4, {"pdj_a", "pdj_b", "pdj_c", "pdj_d"}, 0, 0},
  {"fan2", 2, {"fan2_x", "fan2_y"}, 0, 0},
  {"rings2", 1, {"rings2_val"}, 0, 0},
  {"eyefish", 0, {0}, 0, 0}, {"bubble", 0, {0}, 0, 0},
  {"perspective", 2, {"perspective_angle", "perspective_dist"}, 0, 0},
  {"noise", 0, {0}, 0, 0},
  {"julian", 2, {"julian_power", "julian_dist"}, 0, 0},
  {"juliascope", 2, {"juliascope_power", "juliascope_dist"}, 0, 0},
  {"blur", 0, {0}, 0, 0}, {"gaussian_blur", 0, {0}, 0, 0},
  {"radial_blur", 1, {"radial_blur_angle"}, 1, 0},
  {"pie", 3, {"pie_slices", "pie_rotation", "pie_thickness"}, 0, 0},
  {"ngon", 4, {"ngon_power", "ngon_sides", "ngon_corners", "ngon_circle"}, 0, 0},
  {"curl", 2, {"curl_c1", "curl_c2"}, 0, 0},
  {"rectangles", 2, {"rectangles_x", "rectangles_y"}, 0, 0},
  {"arch", 0, {0}, 0, 0}, {"tangent", 0, {0}, 0, 0}, {"square", 0, {0}, 0, 0}, {"rays", 0, {0}, 0, 0},
  {"blade", 0, {0}, 0, 0}, {"secant2", 0, {0}, 0, 0}, {"twintrian", 0, {0}, 0, 0}, {"cross", 0, {0}, 0, 0},
  {"disc2", 2, {"disc2_rot", "disc2_twist"}, 0, 0},
  {"super_shape", 6, {"super_shape_n1", "super_shape_n2", "super_shape_n3", "super_shape_m", "super_shape_holes",
                      "super_shape_rnd"}, 0, 0},
  {"flower", 2, {"flower_holes", "flower_petals"}, 0, 0},
  {"conic", 2, {"conic_holes", "conic_eccentricity"}, 0, 0},
  {"parabola", 2, {"parabola_height", "parabola_width"}, 0, 0},
  {"bent2", 2, {"bent2_x", "bent2_y"}, 0, 0},
  {"bipolar", 1, {"bipolar_shift"}, 0, 0},
  {"boarders", 0, {0}, 0, 0}, {"butterfly", 0, {0}, 0, 0},
  {"cell", 1, {"cell_size"}, 0, 0},
  {"cpow", 3, {"cpow_power", "cpow_r", "cpow_i"}, 0, 0},
  {"curve", 4, {"curve_xlength", "curve_ylength", "curve_xamp", "curve_yamp"}, 0, 0},
  {"edisc", 0, {0}, 0, 0}, {"foci", 0, {0}, 0, 0},
  {"oscope", 4, {"oscope_separation", "oscope_frequency", "oscope_amplitude", "oscope_damping"}, 0, 0},
  {"loonie", 0, {0}, 0, 0}, {"exp", 0, {0}, 0, 0}, {"log", 0, {0}, 0, 0}, {"sin", 0, {0}, 0, 0},
  {"cos", 0, {0}, 0, 0}, {"sinh", 0, {0}, 0, 0},
  {"pre_blur", 0, {0}, 0, 1},
  {"waves2", 4, {"waves2_scalex", "waves2_scaley", "waves2_freqx", "waves2_freqy"}, 0, 0},
  {"cylinder", 0, {0}, 0, 0}, {"coth", 0, {0}, 0, 0},
  {"auger", 4, {"auger_freq", "auger_scale", "auger_sym", "auger_weight"}, 0, 0},
  {"flux", 1, {"flux_spread"}, 0, 0},
  {"mobius", 8, {"mobius_re_a", "mobius_re_b", "mobius_re_c", "mobius_re_d", "mobius_im_a", "mobius_im_b",
                 "mobius_im_c", "mobius_im_d"}, 0, 0},
};
#define NVARS ((int)(sizeof(VARS) / sizeof(VARS[0])))

int orc_var_count(void) { return NVARS; }
const char* orc_var_name(int id) { return (id >= 0 && id < NVARS) ? VARS[id].name : 0; }
int orc_var_id(const char* name) {
  for (int i = 0; i < NVARS; ++i) if (strcmp(VARS[i].name, name) == 0) return i;
  return -1;
}
int orc_var_nparams(int id) { return (id >= 0 && id < NVARS) ? VARS[id].nprm : -1; }
const char* orc_var_param(int id, int k) {
  return (id >= 0 && id < NVARS && k >= 0 && k < VARS[id].nprm) ? VARS[id].prm[k] : 0;
}

static double mod2(double a, double b) { return a - b * trunc(a / b); }  /* math.glsl mod2 */
static int badval(double v) { return (v != v) || v > 1e10 || v < -1e10; }

/* Each case cites the entry of the reference's variations.yaml it restates.
 * GLSL sincos(t) = (sin t, cos t); `.yx` swaps to (cos t, sin t). */
typedef struct { double rsq, r, a, phi, sina, cosa, sinr, cosr; } precalc;
static void make_precalc(double x, double y, precalc* pc) {
  pc->rsq = x * x + y * y;                        /* :3 */
  pc->r = sqrt(pc->rsq);                          /* :2 length(v) */
  pc->a = atan2(x, y);                            /* :4 atan($x, $y) */
  pc->phi = atan2(y, x);                          /* :5 */
  pc->sina = x / pc->r; pc->cosa = y / pc->r;     /* :6-7 */
  pc->sinr = sin(pc->r); pc->cosr = cos(pc->r);   /* :8-9 */
}

static void var_eval_pc(int id, double x, double y, const precalc* pc, double w, const float* P, const float* A,
                        orc_rng* rng, double* ox, double* oy) {
  const double rsq = pc->rsq, r = pc->r, a = pc->a, phi = pc->phi;
  const double sina = pc->sina, cosa = pc->cosa, sinr = pc->sinr, cosr = pc->cosr;
  const double p0 = P ? P[0] : 0, p1 = P ? P[1] : 0, p2 = P ? P[2] : 0, p3 = P ? P[3] : 0;
  const double p4 = P ? P[4] : 0, p5 = P ? P[5] : 0, p6 = P ? P[6] : 0, p7 = P ? P[7] : 0;
  const double c10 = A[2], c11 = A[3], c20 = A[4], c21 = A[5];
  double X = 0, Y = 0;
  switch (id) {
    case 0: X = x; Y = y; break;                                              /* linear :13 */
    case 1: X = sin(x); Y = sin(y); break;                                    /* sinusoidal :16 */
    case 2: X = x / (rsq + EPS_D); Y = y / (rsq + EPS_D); break;              /* spherical :19 */
    case 3: { double c1 = sin(rsq), c2 = cos(rsq);                            /* swirl :22 */
      X = c1 * x - c2 * y; Y = c2 * x + c1 * y; } break;
    case 4: X = (x + y) * (x - y) / (r + EPS_D); Y = 2.0 * x * y / (r + EPS_D); break;  /* horseshoe :28 */
    case 5: X = a / PI_D; Y = r - 1.0; break;                                 /* polar :31 */
    case 6: X = r * sin(a + r); Y = r * cos(a - r); break;                    /* handkerchief :34 */
    case 7: X = r * sin(r * a); Y = -r * cos(r * a); break;                   /* heart :37 */
    case 8: X = a * sin(r * PI_D) / PI_D; Y = a * cos(r * PI_D) / PI_D; break; /* disc :40 */
    case 9: X = (cosa + sinr) / (r + EPS_D); Y = (sina - cosr) / (r + EPS_D); break; /* spiral :43 */
    case 10: X = sina / (r + EPS_D); Y = cosa * r; break;                     /* hyperbolic :46 */
    case 11: X = sina * cosr; Y = cosa * sinr; break;                         /* diamond :49 */
    case 12: { double n0 = sin(a + r), n1 = cos(a - r);                       /* ex :52 */
      double m0 = n0 * n0 * n0 * r, m1 = n1 * n1 * n1 * r; X = m0 + m1; Y = m0 - m1; } break;
    case 13: { double t = 0.5 * a + nearbyint(orc_rng_next(rng)) * PI_D;     /* julia :60 */
      X = sqrt(r) * cos(t); Y = sqrt(r) * sin(t); } break;
    case 14: X = x < 0 ? 2.0 * x : x; Y = y < 0 ? y / 2.0 : y; break;         /* bent :64 */
    case 15: X = x + c10 * sin(y / (c20 * c20 + EPS_D));                      /* waves :71 */
      Y = y + c11 * sin(x / (c21 * c21 + EPS_D)); break;
    case 16: X = 2.0 / (r + 1.0) * y; Y = 2.0 / (r + 1.0) * x; break;         /* fisheye :78 */
    case 17: X = x + c20 * sin(tan(3.0 * y)); Y = y + c21 * sin(tan(3.0 * x)); break; /* popcorn :81 */
    case 18: X = exp(x - 1.0) * cos(PI_D * y); Y = exp(x - 1.0) * sin(PI_D * y); break; /* exponential :88 */
    case 19: X = pow(r, sina) * cosa; Y = pow(r, sina) * sina; break;         /* power :91 */
    case 20: X = cos(PI_D * x) * cosh(y); Y = -sin(PI_D * x) * sinh(y); break; /* cosine :94 */
    case 21: { double dx = c20 * c20 + EPS_D;                                 /* rings :97 */
      double m = mod2(r + dx, 2.0 * dx) - dx + r * (1.0 - dx); X = m * cosa; Y = m * sina; } break;
    case 22: { double dx = PI_D * (c20 * c20 + EPS_D), dx2 = dx * 0.5, t = a; /* fan :102 */
      t += (mod2(t + c21, dx) > dx2) ? -dx2 : dx2; X = r * cos(t); Y = r * sin(t); } break;
    case 23: { double m = r * (p1 + (p0 - p1) * (0.5 + 0.5 * sin(p2 * a)));   /* blob :110 */
      X = m * sina; Y = m * cosa; } break;
    case 24: X = sin(p0 * y) - cos(p1 * x); Y = sin(p2 * x) - cos(p3 * y); break; /* pdj :117 */
    case 25: { double dx = PI_D * (p0 * p0 + EPS_D), dx2 = 0.5 * dx;           /* fan2 :129 */
      double t = a + p1 - dx * floor((a + p1) / dx); double fa = t > dx2 ? a - dx2 : a + dx2;
      X = r * sin(fa); Y = r * cos(fa); } break;
    case 26: { double dx = p0 * p0 + EPS_D, rr = r;                           /* rings2 :143 */
      rr += -2.0 * dx * (double)(int)((rr + dx) / (2.0 * dx)) + rr * (1.0 - dx); X = rr * sina; Y = rr * cosa; } break;
    case 27: X = 2.0 / (r + 1.0) * x; Y = 2.0 / (r + 1.0) * y; break;         /* eyefish :152 */
    case 28: X = x / (rsq * 0.25 + 1.0); Y = y / (rsq * 0.25 + 1.0); break;   /* bubble :155 */
    case 29: { double d = p1 - y * sin(p0);                                    /* perspective :158 */
      X = x * p1 / d; Y = p1 * cos(p0) * y / d; } break;
    case 30: { double m = orc_rng_next(rng); double t = orc_rng_next(rng) * 2.0 * PI_D; /* noise :166 */
      X = x * m * cos(t); Y = y * m * sin(t); } break;
    case 31: { int k = (int)trunc(fabs(p0) * orc_rng_next(rng));              /* julian :169 */
      double m = pow(r, p1 / p0), t = (phi + 2.0 * PI_D * k) / p0; X = m * cos(t); Y = m * sin(t); } break;
    case 32: { int k = (int)trunc(fabs(p0) * orc_rng_next(rng));              /* juliascope :182 */
      double m = pow(r, p1 / p0), t = (2.0 * PI_D * k + ((k & 1) == 0 ? 1.0 : -1.0) * phi) / p0;
      X = m * cos(t); Y = m * sin(t); } break;
    case 33: { double m = orc_rng_next(rng); double t = orc_rng_next(rng) * 2.0 * PI_D; /* blur :195 */
      X = m * cos(t); Y = m * sin(t); } break;
    case 34: { double g = orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng);
      g -= 2.0; double t = orc_rng_next(rng) * 2.0 * PI_D; X = g * cos(t); Y = g * sin(t); } break; /* gaussian_blur :198 */
    case 35: { double sp = sin(p0 * PI_D / 2.0), zm = cos(p0 * PI_D / 2.0);    /* radial_blur :201 */
      double g = orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng);
      g = w * (g - 2.0); double t = phi + sp * g, rz = zm * g - 1.0;
      X = r * cos(t) + rz * x; Y = r * sin(t) + rz * y; } break;
    case 36: { double m = orc_rng_next(rng); double sl = trunc(orc_rng_next(rng) * p0 + 0.5); /* pie :212 */
      double t = p1 + 2.0 * PI_D * (sl + orc_rng_next(rng) * p2) / p0; X = m * sin(t); Y = m * cos(t); } break;
    case 37: { double rf = pow(rsq, p0 * 0.5), b = 2.0 * PI_D / (p1 + EPS_D); /* ngon :219 */
      double ph = phi - b * floor(a / b); if (ph > b * 0.5) ph -= b;
      double m = (p2 * (1.0 / (cos(ph) + EPS_D) - 1.0) + p3) / (rf + EPS_D); X = m * x; Y = m * y; } break;
    case 38: { double re = 1.0 + p0 * x + p1 * (x * x - y * y), im = p0 * y + 2.0 * p1 * x * y; /* curl :232 */
      double d = re * re + im * im; X = (x * re + y * im) / d; Y = (y * re - x * im) / d; } break;
    case 39: X = p0 == 0 ? x : (2.0 * floor(x / p0) + 1.0) * p0 - x;          /* rectangles :241 */
      Y = p1 == 0 ? y : (2.0 * floor(y / p1) + 1.0) * p1 - y; break;
    case 40: { double t = orc_rng_next(rng) * w * PI_D, s = sin(t), c = cos(t); X = s; Y = s * s / c; } break; /* arch :251 */
    case 41: X = sin(x) / cos(y); Y = tan(y); break;                          /* tangent :256 */
    case 42: X = orc_rng_next(rng) - 0.5; Y = orc_rng_next(rng) - 0.5; break; /* square :259 */
    case 43: { double ang = w * orc_rng_next(rng) * PI_D, rr = w / (rsq + EPS_D); /* rays :262 */
      X = tan(ang) * rr * cos(x); Y = tan(ang) * rr * sin(y); } break;
    case 44: { double t = orc_rng_next(rng) * w * r, s = sin(t), c = cos(t);  /* blade :268 */
      X = x * (c + s); Y = x * (c - s); } break;
    case 45: { double cr = cos(w * r); X = x; Y = 1.0 / cr + (cr < 0 ? 1.0 : -1.0); } break; /* secant2 :273 */
    case 46: { double t = orc_rng_next(rng) * w * r, s = sin(t), c = cos(t);  /* twintrian :279 (flam3 form) */
      double diff = log10(s * s) + c; if (badval(diff)) diff = -30.0; X = x * diff; Y = x * (diff - s * PI_D); } break;
    case 47: { double s = x * x - y * y, m = sqrt(1.0 / (s * s + EPS_D)); X = x * m; Y = y * m; } break; /* cross :286 */
    case 48: { double tp = p0 * PI_D, sa = sin(p1), ca = cos(p1) - 1.0;       /* disc2 :291 */
      double k = 1.0 + p1 + (p1 > 2.0 * PI_D ? -2.0 : 2.0) * PI_D; sa *= k; ca *= k;
      double t = tp * (x + y); X = (sin(t) + ca) * a / PI_D; Y = (cos(t) + sa) * a / PI_D; } break;
    case 49: { double t = p3 * 0.25 * phi + PI_D * 0.25, s = sin(t), c = cos(t); /* super_shape :303 */
      double mixed = r * (1.0 - p5) + orc_rng_next(rng) * p5;
      double m = (mixed - p4) * pow(pow(fabs(c), p1) + pow(fabs(s), p2), -1.0 / p0) / r; X = x * m; Y = y * m; } break;
    case 50: { double m = (orc_rng_next(rng) - p0) * cos(p1 * phi) / r; X = x * m; Y = y * m; } break; /* flower :315 */
    case 51: { double m = (orc_rng_next(rng) - p0) * p1 / (1.0 + p1 * x / r) / r; X = x * m; Y = y * m; } break; /* conic :321 */
    case 52: X = p0 * sinr * sinr * orc_rng_next(rng); Y = p1 * cosr * orc_rng_next(rng); break; /* parabola :327 */
    case 53: X = x < 0 ? x * p0 : x; Y = y < 0 ? y * p1 : y; break;           /* bent2 :333 (flam3 form) */
    case 54: { double t = rsq + 1.0, x2 = 2.0 * x, ps = -(PI_D / 2.0) * p0;   /* bipolar :339 (flam3 form) */
      double yy = 0.5 * atan2(2.0 * y, rsq - 1.0) + ps;
      if (yy > PI_D / 2.0) yy = -PI_D / 2.0 + mod2(yy + PI_D / 2.0, PI_D);
      else if (yy < -PI_D / 2.0) yy = PI_D / 2.0 - mod2(PI_D / 2.0 - yy, PI_D);
      X = 0.25 * (2.0 / PI_D) * log((t + x2) / (t - x2)); Y = (2.0 / PI_D) * yy; } break;
    case 55: { double rx = nearbyint(x), ry = nearbyint(y), dx = x - rx, dy = y - ry; /* boarders :351 */
      if (orc_rng_next(rng) >= 0.75) { X = dx * 0.5 + rx; Y = dy * 0.5 + ry; }
      else if (fabs(dx) >= fabs(dy)) {
        if (dx > 0) { X = dx * 0.5 + rx + 0.25; Y = dy * 0.5 + ry + 0.25 * dy / dx; }
        else { X = dx * 0.5 + rx - 0.25; Y = dy * 0.5 + ry - 0.25 * dy / dx; }
      } else {
        if (dy > 0) { X = dx * 0.5 + rx + 0.25 * dx / dy; Y = dy * 0.5 + ry + 0.25; }
        else { X = dx * 0.5 + rx - 0.25 * dx / dy; Y = dy * 0.5 + ry - 0.25; }
      } } break;
    case 56: { double y2 = 2.0 * y;                                           /* butterfly :392 */
      double m = 1.3029400317411197908970256609023 * sqrt(fabs(x * y) / (EPS_D + x * x + y2 * y2)); X = m * x; Y = m * y2; } break;
    case 57: { double inv = 1.0 / p0; int cx = (int)floor(x * inv), cy = (int)floor(y * inv); /* cell :397 (flam3 form) */
      double dx = x - cx * p0, dy = y - cy * p0;
      if (cy >= 0) { if (cx >= 0) { cx *= 2; cy *= 2; } else { cx = -(2 * cx + 1); cy *= 2; } }
      else { if (cx >= 0) { cx *= 2; cy = -(2 * cy + 1); } else { cx = -(2 * cx + 1); cy = -(2 * cy + 1); } }
      X = dx + cx * p0; Y = dy + cy * p0; } break;
    case 58: { double lnr = 0.5 * log(rsq), va = 2.0 * PI_D / p0, vc = p1 / p0, vd = p2 / p0; /* cpow :413 */
      double m = exp(vc * lnr - vd * phi), t = vc * phi + vd * lnr + va * floor(p0 * orc_rng_next(rng));
      X = m * cos(t); Y = m * sin(t); } break;
    case 59: X = x + p2 * exp(-y * y / fmax(p0 * p0, 1e-20));                  /* curve :425 */
      Y = y + p3 * exp(-x * x / fmax(p1 * p1, 1e-20)); break;
    case 60: { double t = rsq + 1.0, t2 = 2.0 * x;                            /* edisc :437 (flam3 form) */
      double xmax = (sqrt(t + t2) + sqrt(t - t2)) * 0.5, a1 = log(xmax + sqrt(xmax - 1.0)), a2 = -acos(x / xmax);
      double sv = sin(a1), cv = cos(a1); if (y > 0) sv = -sv;
      X = cosh(a2) * cv / 11.57034632; Y = sinh(a2) * sv / 11.57034632; } break;
    case 61: { double ex = exp(x) * 0.5, enx = 0.25 / ex, d = ex + enx - cos(y); /* foci :450 */
      X = (ex - enx) / d; Y = sin(y) / d; } break;
    case 62: { double tpf = 2.0 * PI_D * p1;                                  /* oscope :457 (flam3 form) */
      double t = p3 == 0 ? p2 * cos(tpf * x) + p0 : p2 * exp(-fabs(x) * p3) * cos(tpf * x) + p0;
      X = x; Y = fabs(y) <= t ? -y : y; } break;
    case 63: { double w2 = w * w, m = rsq < w2 ? sqrt(w2 / rsq - 1.0) : 1.0; X = x * m; Y = y * m; } break; /* loonie :459 */
    case 64: X = exp(x) * cos(y); Y = exp(x) * sin(y); break;                 /* exp :464 */
    case 65: X = 0.5 * log(rsq); Y = phi; break;                              /* log :467 */
    case 66: X = sin(x) * cosh(y); Y = cos(x) * sinh(y); break;               /* sin :470 */
    case 67: X = cos(x) * cosh(y); Y = sin(x) * sinh(y); break;               /* cos :473 */
    case 68: X = sinh(x) * cos(y); Y = cosh(x) * sin(y); break;               /* sinh :476 */
    case 69: { double g = orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng); g += orc_rng_next(rng);
      g -= 2.0; double t = orc_rng_next(rng) * 2.0 * PI_D; X = g * cos(t); Y = g * sin(t); } break; /* pre_blur :481 */
    case 70: X = x + p0 * sin(y * p2); Y = y + p1 * sin(x * p3); break;       /* waves2 :485 */
    case 71: X = sin(x); Y = y; break;                                        /* cylinder :497 */
    case 72: { double den = 1.0 / (cosh(2.0 * x) - cos(2.0 * y)); X = den * sinh(2.0 * x); Y = den * sin(2.0 * y); } break; /* coth :500 */
    case 73: { double s = sin(x * p0), t = sin(y * p0);                       /* auger :508 */
      double dy = y + p3 * (p1 * s / 2.0 + fabs(y) * s), dx = x + p3 * (p1 * t / 2.0 + fabs(x) * t);
      X = x + p2 * (dx - x); Y = dy; } break;
    case 74: { double xpw = x + w, xmw = x - w;                               /* flux :521 */
      double m = (2.0 + p0) * sqrt(sqrt(y * y + xpw * xpw) / sqrt(y * y + xmw * xmw));
      double t = (atan2(y, xmw) - atan2(y, xpw)) * 0.5; X = m * cos(t); Y = m * sin(t); } break;
    case 75: { double ure = p0 * x - p4 * y + p1, uim = p0 * y + p4 * x + p5; /* mobius :532 */
      double vre = p2 * x - p6 * y + p3, vim = p2 * y + p6 * x + p7, d = vre * vre + vim * vim;
      X = (ure * vre + uim * vim) / d; Y = (uim * vre - ure * vim) / d; } break;
    default: X = 0; Y = 0; break;
  }
  *ox = X;
  *oy = Y;
}

void orc_var_eval(int id, double x, double y, double w, const float* P, const float* A, orc_rng* rng,
                  double* ox, double* oy) {
  precalc pc;
  make_precalc(x, y, &pc);
  var_eval_pc(id, x, y, &pc, w, P, A, rng, ox, oy);
}

/* ==================================================================== xforms */
/* Terms in std::map order (src/flame.hpp:26): one per name, last wins,
 * name-sorted; pre_xform terms first (variation_table.cpp:115-117). */
static int sorted_terms(const orc_xform* xf, int* idx) {
  int n = 0;
  for (int t = 0; t < xf->n_terms; ++t) {
    int dup = 0;
    for (int q = t + 1; q < xf->n_terms; ++q) if (xf->terms[q].var == xf->terms[t].var) dup = 1;
    if (!dup && xf->terms[t].var >= 0 && xf->terms[t].var < NVARS) idx[n++] = t;
  }
  for (int i = 1; i < n; ++i) {  /* insertion sort: pre flag, then name */
    int v = idx[i], j = i - 1;
    while (j >= 0) {
      const var_def* A = &VARS[xf->terms[idx[j]].var];
      const var_def* B = &VARS[xf->terms[v].var];
      int gt = (A->pre_xform != B->pre_xform) ? (A->pre_xform < B->pre_xform) : (strcmp(A->name, B->name) > 0);
      if (!gt) break;
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
  return n;
}

/* A copy of the xform with its terms deduplicated and in evaluation order. */
static void prepare_xform(const orc_xform* src, orc_xform* dst) {
  int idx[ORC_MAX_TERMS];
  int n = sorted_terms(src, idx);
  *dst = *src;
  for (int k = 0; k < n; ++k) dst->terms[k] = src->terms[idx[k]];
  dst->n_terms = n;
}
static void prepare_flame(const orc_flame* src, orc_flame* dst) {
  *dst = *src;
  for (int i = 0; i < src->n_xforms; ++i) prepare_xform(&src->xf[i], &dst->xf[i]);
  if (src->has_final) prepare_xform(&src->fin, &dst->fin);
}

static float mixf(float a, float b, float t) { float t0 = a * (1.0f - t); float t1 = b * t; return t0 + t1; }

static void apply_f32_sorted(const orc_xform* xf, float x0, float y0, float c0, int first, orc_rng* rng,
                             float out[4]) {
  const float* A = xf->aff;
  float x = fmaf(A[0], x0, fmaf(A[2], y0, A[4]));  /* variation_table.cpp:105 */
  float y = fmaf(A[1], x0, fmaf(A[3], y0, A[5]));
  const int n = xf->n_terms;
  float rx = 0.0f, ry = 0.0f;
  int first_reg = 1;
  int have_pc = 0;
  precalc pc;
  for (int k = 0; k < n; ++k) {
    const orc_term* t = &xf->terms[k];
    const var_def* v = &VARS[t->var];
    double dx, dy;
    if (v->pre_xform || !have_pc) { make_precalc((double)x, (double)y, &pc); have_pc = !v->pre_xform; }
    var_eval_pc(t->var, (double)x, (double)y, &pc, (double)t->weight, t->prm, A, rng, &dx, &dy);
    float fx = (float)dx, fy = (float)dy;
    if (!v->no_weight_mul) { fx = t->weight * fx; fy = t->weight * fy; }
    if (v->pre_xform) { x = x + fx; y = y + fy; continue; }
    if (first_reg) { rx = fx; ry = fy; first_reg = 0; }
    else { rx = rx + fx; ry = ry + fy; }
  }
  if (xf->has_post) {                              /* variation_table.cpp:154-158 */
    const float* Q = xf->post;
    float px = fmaf(Q[0], rx, fmaf(Q[2], ry, Q[4]));
    float py = fmaf(Q[1], rx, fmaf(Q[3], ry, Q[5]));
    rx = px; ry = py;
  }
  float cin = first ? (float)orc_rng_next(rng) : c0;  /* variation_table.cpp:241-243 */
  out[0] = rx; out[1] = ry;
  out[2] = mixf(cin, xf->color, xf->color_speed);
  out[3] = xf->opacity;
}

void orc_apply_f32(const orc_xform* xf, float x0, float y0, float c0, int first, orc_rng* rng, float out[4]) {
  orc_xform p;
  prepare_xform(xf, &p);
  apply_f32_sorted(&p, x0, y0, c0, first, rng, out);
}

static void apply_f64_sorted(const orc_xform* xf, double x0, double y0, double c0, int first, orc_rng* rng,
                             double out[4]) {
  const float* A = xf->aff;
  double x = fma(A[0], x0, fma(A[2], y0, A[4]));
  double y = fma(A[1], x0, fma(A[3], y0, A[5]));
  const int n = xf->n_terms;
  double rx = 0.0, ry = 0.0;
  int have_pc = 0;
  precalc pc;
  for (int k = 0; k < n; ++k) {
    const orc_term* t = &xf->terms[k];
    const var_def* v = &VARS[t->var];
    double dx, dy;
    if (v->pre_xform || !have_pc) { make_precalc(x, y, &pc); have_pc = !v->pre_xform; }
    var_eval_pc(t->var, x, y, &pc, (double)t->weight, t->prm, A, rng, &dx, &dy);
    if (!v->no_weight_mul) { dx *= t->weight; dy *= t->weight; }
    if (v->pre_xform) { x += dx; y += dy; continue; }
    rx += dx; ry += dy;
  }
  if (xf->has_post) {
    const float* Q = xf->post;
    double px = fma(Q[0], rx, fma(Q[2], ry, Q[4]));
    double py = fma(Q[1], rx, fma(Q[3], ry, Q[5]));
    rx = px; ry = py;
  }
  double cin = first ? orc_rng_next(rng) : c0;
  out[0] = rx; out[1] = ry;
  out[2] = cin * (1.0 - xf->color_speed) + xf->color * xf->color_speed;
  out[3] = xf->opacity;
}

void orc_apply_f64(const orc_xform* xf, double x0, double y0, double c0, int first, orc_rng* rng, double out[4]) {
  orc_xform p;
  prepare_xform(xf, &p);
  apply_f64_sorted(&p, x0, y0, c0, first, rng, out);
}

void orc_cdf(const orc_flame* f, float cdf[ORC_MAX_XF]) {
  float total = 0.0f;                              /* flame.cpp:78-82 */
  for (int i = 0; i < f->n_xforms; ++i) total += f->xf[i].weight;
  float sum = 0.0f;
  for (int i = 0; i < f->n_xforms; ++i) {
    float wn = f->xf[i].weight / total;
    sum = (i == 0) ? wn : sum + wn;
    cdf[i] = sum;
  }
}

int orc_select(const orc_flame* f, float u) {      /* xform_select.tpl.glsl:1-14 */
  float cdf[ORC_MAX_XF];
  orc_cdf(f, cdf);
  for (int i = 0; i < f->n_xforms - 1; ++i) if (cdf[i] >= u) return i;
  return f->n_xforms - 1;
}

void orc_animate(const orc_flame* src, uint32_t i, uint32_t T, float tss, orc_flame* dst) {
  *dst = *src;                                     /* animate.tpl.glsl:18-96 */
  /* half_width = T/2 (animate.tpl.glsl:19, T/32 groups of 32 lanes,
   * flame.cpp:248); the reference leaves T < 32 undefined, here T = 1 is the
   * centre sample (dt = 0).  sin/cos are evaluated in double and rounded, the
   * division is correctly rounded (the GPU kernel does the same). */
  const int half = (int)(T / 2u);
  const int pos = (int)i - half;                   /* animate.tpl.glsl:20 */
  const float dt = half ? (float)pos / (float)half * tss : 0.0f;   /* :21 */
  const float rps = 0.31415926535f;                /* :16 */
  for (int k = 0; k <= src->n_xforms; ++k) {
    const orc_xform* s;
    orc_xform* d;
    if (k < src->n_xforms) { s = &src->xf[k]; d = &dst->xf[k]; }
    else if (src->has_final) { s = &src->fin; d = &dst->fin; }
    else break;
    float th = rps * dt * s->rotation_frequency;   /* :43-44 */
    float sn = (float)sin((double)th), cs = (float)cos((double)th);
    const float* a = s->aff;                       /* :45-48 */
    float t0 = a[0] * cs, t1 = a[2] * sn, t2 = a[1] * cs, t3 = a[3] * sn;
    float t4 = a[2] * cs, t5 = a[0] * sn, t6 = a[3] * cs, t7 = a[1] * sn;
    d->aff[0] = t0 + t1; d->aff[1] = t2 + t3; d->aff[2] = t4 - t5; d->aff[3] = t6 - t7;
  }
}

void orc_screen_affine(uint32_t W, uint32_t H, const uint32_t size[2], const float center[2], float scale,
                       float rotate, float out[6]) {
  float b[6] = {1, 0, 0, 1, 0, 0};                 /* flame.cpp:289-296, flame.hpp:97-128 */
  float t0 = (float)W / 2.0f, t1 = (float)H / 2.0f;
  float r4 = b[0] * t0 + b[2] * t1 + b[4], r5 = b[1] * t0 + b[3] * t1 + b[5];
  b[4] = r4; b[5] = r5;
  float s = scale * (float)H / (float)size[1];
  for (int k = 0; k < 4; ++k) b[k] = b[k] * s;
  float rad = 0.01745329251f * rotate;
  float sino = sinf(rad), coso = cosf(rad);
  float q0 = b[0] * coso + b[2] * sino, q1 = b[1] * coso + b[3] * sino;
  float q2 = b[2] * coso - b[0] * sino, q3 = b[3] * coso - b[1] * sino;
  b[0] = q0; b[1] = q1; b[2] = q2; b[3] = q3;
  t0 = -center[0]; t1 = -center[1];
  r4 = b[0] * t0 + b[2] * t1 + b[4]; r5 = b[1] * t0 + b[3] * t1 + b[5];
  b[4] = r4; b[5] = r5;
  for (int k = 0; k < 6; ++k) out[k] = b[k];
}

static uint32_t rev32(uint32_t n) {                /* hammersley.cpp:5-11 */
  n = (n << 16) | (n >> 16);
  n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
  n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
  n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
  n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
  return n;
}

void orc_hammersley(uint32_t count, float* out) {  /* hammersley.cpp:29-48 */
  uint32_t mx = count;
  if (count % 2 != 0) { uint32_t v = count - 1; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; mx = v + 1; }
  float inv = 1.0f / (float)mx;
  uint32_t r = 0, v = mx;
  while (v >>= 1) ++r;
  for (uint32_t i = 0; i < count; ++i) {
    uint32_t fl = r ? rev32(i) >> (32 - r) : 0;
    out[4 * (size_t)i + 0] = (float)((double)((float)i * inv) * 2.0 - 1.0);
    out[4 * (size_t)i + 1] = (float)((double)((float)fl * inv) * 2.0 - 1.0);
    out[4 * (size_t)i + 2] = 0.0f;
    out[4 * (size_t)i + 3] = 0.0f;
  }
}

int orc_pal_index(float c) {                       /* flame.glsl:84 */
  float t = ceilf(c * 255.0f);
  if (!(t > 0.0f)) return 0;
  if (t >= 255.0f) return 255;
  return (int)t;
}

/* Taylor-series sin/cos on [0,1] with fmaf, the same evaluation as the
 * library uses for the first-run jitter angle. */
static void sincos01(float q, float* s, float* c) {
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

/* flame.glsl:74-87: final xform (drawing only), screen fma affine, floor,
 * bounds and w > 0, bin (H-1-y)*W + x. */
static int draw_f32(const orc_flame* f, const float o[4], orc_rng* rng, uint32_t W, uint32_t H, const float ss[6],
                    double* bins, uint32_t* hits) {
  float rx = o[0], ry = o[1], rc = o[2], rw = o[3];
  if (f->has_final) {
    float q[4];
    apply_f32_sorted(&f->fin, o[0], o[1], o[2], 0, rng, q);
    rx = q[0]; ry = q[1]; rc = q[2]; rw = q[3] * o[3];
  }
  float sx = fmaf(ss[0], rx, fmaf(ss[2], ry, ss[4]));
  float sy = fmaf(ss[1], rx, fmaf(ss[3], ry, ss[5]));
  float fx = floorf(sx), fy = floorf(sy);
  if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)W && fy < (float)H && rw > 0.0f)) return 0;
  size_t b = (size_t)(H - (uint32_t)fy - 1u) * W + (uint32_t)fx;
  int pi = orc_pal_index(rc);
  if (bins) {
    bins[4 * b + 0] += f->pal[pi][0];
    bins[4 * b + 1] += f->pal[pi][1];
    bins[4 * b + 2] += f->pal[pi][2];
    bins[4 * b + 3] += rw;
  }
  if (hits) hits[b] += 1;
  return 1;
}

void orc_compat_pass(const orc_flame* flames, int stride, uint32_t T, uint32_t ppt, const float* pos_in,
                     float* pos_out, const uint32_t* shuf, uint32_t* rs, uint32_t in_row, uint32_t out_row,
                     int random_read, int random_write, int first_run, int do_draw, uint32_t W, uint32_t H,
                     const float ss[6], double* bins, uint32_t* hits, uint64_t* drawn) {
  uint64_t cnt = 0;
  orc_flame* prep = (orc_flame*)malloc(sizeof(orc_flame));
  int prepared_for = -1;
  for (uint32_t y = 0; y < T; ++y) {
    const int fi = stride ? (int)y : 0;
    if (fi != prepared_for) { prepare_flame(flames + fi, prep); prepared_for = fi; }
    const orc_flame* f = prep;
    for (uint32_t wg = 0; wg < ppt / 256u; ++wg) {
      int xid = 0;
      for (uint32_t l = 0; l < 256u; ++l) {
        const uint32_t gx = wg * 256u + l;
        const size_t slot = (size_t)y * ppt + gx;
        orc_jsf st = {rs[4 * slot], rs[4 * slot + 1], rs[4 * slot + 2], rs[4 * slot + 3]};
        orc_rng rng = {0, &st, 0, 0};
        if (l == 0) xid = orc_select(f, orc_jsf_randf(&st));       /* flame.glsl:51-53 */
        uint32_t i_idx = random_read ? shuf[(size_t)in_row * ppt + gx] : gx;
        uint32_t o_idx = random_write ? shuf[(size_t)out_row * ppt + gx] : gx;
        size_t off = first_run ? 0 : (size_t)y * ppt;
        const float* p = pos_in + 4 * (off + i_idx);
        float px = p[0], py = p[1], pc = p[2];
        if (first_run) {                                              /* flame.glsl:63-66 */
          float r1 = orc_jsf_randf(&st), r2 = orc_jsf_randf(&st);
          float s = ((r1 * 0.1f) * PI_F) * 2.0f, sn, cs;
          sincos01(sqrtf(r2), &sn, &cs);
          px = px + s * sn;
          py = py + s * cs;
        }
        float o[4];
        apply_f32_sorted(&f->xf[xid], px, py, pc, first_run, &rng, o);
        float* q = pos_out + 4 * ((size_t)y * ppt + o_idx);
        q[0] = o[0]; q[1] = o[1]; q[2] = o[2]; q[3] = 0.0f;
        if (do_draw) cnt += (uint64_t)draw_f32(f, o, &rng, W, H, ss, bins, hits);
        rs[4 * slot] = st.a; rs[4 * slot + 1] = st.b; rs[4 * slot + 2] = st.c; rs[4 * slot + 3] = st.d;
      }
    }
  }
  free(prep);
  if (drawn) *drawn = cnt;
}

void orc_jsf_seed_all(uint32_t* out, uint64_t n) {   /* flame.cpp:139-141 for slots 0..n-1 */
  for (uint64_t i = 0; i < n; ++i) {
    orc_jsf s;
    orc_jsf_seed(&s, (uint32_t)i);
    out[4 * i] = s.a; out[4 * i + 1] = s.b; out[4 * i + 2] = s.c; out[4 * i + 3] = s.d;
  }
}

/* ---------------------------------------- reference passes, multithreaded */
/* The reference's passes (flame.glsl:41-90 as orc_compat_pass restates it),
 * spread over threads by workgroup, with a float RGBA histogram that every
 * thread adds to without atomics as the shader does (flame.glsl:82-84: a
 * read-modify-write that can lose updates when two invocations hit one bin;
 * relaxed atomic loads and stores keep that race defined in C).  Passes are
 * separated by a barrier (the glMemoryBarrier after each dispatch,
 * flame.cpp:262-278, 304-326).  Used as the CPU baseline of the benchmark. */
typedef struct ref_mt_shared {
  const orc_flame* f;
  uint32_t T, ppt, n_passes, W, H;
  float* pos[2];
  const uint32_t* shuf;
  uint32_t* rs;
  const uint32_t* rows;
  int random_read, random_write, first_run, do_draw;
  const float* ss;
  float* bins;
  int threads;
  pthread_barrier_t bar;
} ref_mt_shared;

typedef struct ref_mt_job { ref_mt_shared* sh; int t; uint64_t drawn; } ref_mt_job;

static void add_relaxed(float* p, float v) {
  float cur;
  __atomic_load(p, &cur, __ATOMIC_RELAXED);
  cur += v;
  __atomic_store(p, &cur, __ATOMIC_RELAXED);
}

static void* ref_mt_worker(void* arg) {
  ref_mt_job* J = (ref_mt_job*)arg;
  ref_mt_shared* S = J->sh;
  orc_flame* prep = (orc_flame*)malloc(sizeof(orc_flame));
  prepare_flame(S->f, prep);
  const orc_flame* f = prep;
  const uint32_t groups = S->ppt / 256u, total = S->T * groups;
  const uint32_t g0 = (uint32_t)((uint64_t)total * J->t / S->threads);
  const uint32_t g1 = (uint32_t)((uint64_t)total * (J->t + 1) / S->threads);
  uint64_t cnt = 0;
  for (uint32_t pass = 0; pass < S->n_passes; ++pass) {
    const float* pos_in = S->pos[pass % 2];
    float* pos_out = S->pos[(pass + 1) % 2];
    const uint32_t in_row = S->rows[2 * pass], out_row = S->rows[2 * pass + 1];
    const int first = S->first_run && pass == 0;
    for (uint32_t g = g0; g < g1; ++g) {
      const uint32_t y = g / groups, wg = g % groups;
      int xid = 0;
      for (uint32_t l = 0; l < 256u; ++l) {
        const uint32_t gx = wg * 256u + l;
        const size_t slot = (size_t)y * S->ppt + gx;
        uint32_t* r = S->rs + 4 * slot;
        orc_jsf st = {r[0], r[1], r[2], r[3]};
        orc_rng rng = {0, &st, 0, 0};
        if (l == 0) xid = orc_select(f, orc_jsf_randf(&st));
        const uint32_t i_idx = S->random_read ? S->shuf[(size_t)in_row * S->ppt + gx] : gx;
        const uint32_t o_idx = S->random_write ? S->shuf[(size_t)out_row * S->ppt + gx] : gx;
        const float* p = pos_in + 4 * ((first ? 0 : (size_t)y * S->ppt) + i_idx);
        float px = p[0], py = p[1], pc = p[2];
        if (first) {
          float r1 = orc_jsf_randf(&st), r2 = orc_jsf_randf(&st);
          float sj = ((r1 * 0.1f) * PI_F) * 2.0f, sn, cs;
          sincos01(sqrtf(r2), &sn, &cs);
          px = px + sj * sn;
          py = py + sj * cs;
        }
        float o[4];
        apply_f32_sorted(&f->xf[xid], px, py, pc, first, &rng, o);
        float* q = pos_out + 4 * ((size_t)y * S->ppt + o_idx);
        q[0] = o[0]; q[1] = o[1]; q[2] = o[2]; q[3] = 0.0f;
        if (S->do_draw) {
          float rx = o[0], ry = o[1], rc = o[2], rw = o[3];
          if (f->has_final) {
            float fq[4];
            apply_f32_sorted(&f->fin, o[0], o[1], o[2], 0, &rng, fq);
            rx = fq[0]; ry = fq[1]; rc = fq[2]; rw = fq[3] * o[3];
          }
          const float sx = fmaf(S->ss[0], rx, fmaf(S->ss[2], ry, S->ss[4]));
          const float sy = fmaf(S->ss[1], rx, fmaf(S->ss[3], ry, S->ss[5]));
          const float fx = floorf(sx), fy = floorf(sy);
          if (fx >= 0.0f && fy >= 0.0f && fx < (float)S->W && fy < (float)S->H && rw > 0.0f) {
            const size_t b = (size_t)(S->H - (uint32_t)fy - 1u) * S->W + (uint32_t)fx;
            const int pi = orc_pal_index(rc);
            add_relaxed(S->bins + 4 * b + 0, f->pal[pi][0]);
            add_relaxed(S->bins + 4 * b + 1, f->pal[pi][1]);
            add_relaxed(S->bins + 4 * b + 2, f->pal[pi][2]);
            add_relaxed(S->bins + 4 * b + 3, rw);
            ++cnt;
          }
        }
        r[0] = st.a; r[1] = st.b; r[2] = st.c; r[3] = st.d;
      }
    }
    pthread_barrier_wait(&S->bar);
  }
  free(prep);
  J->drawn = cnt;
  return 0;
}

uint64_t orc_ref_passes_mt(const orc_flame* f, uint32_t T, uint32_t ppt, float* pos_a, float* pos_b,
                           const uint32_t* shuf, uint32_t* rs, const uint32_t* rows, uint32_t n_passes,
                           int random_read, int random_write, int first_run, int do_draw, uint32_t W, uint32_t H,
                           const float ss[6], float* bins, int threads) {
  if (threads < 1) threads = 1;
  ref_mt_shared S;
  S.f = f; S.T = T; S.ppt = ppt; S.n_passes = n_passes; S.W = W; S.H = H;
  S.pos[0] = pos_a; S.pos[1] = pos_b;
  S.shuf = shuf; S.rs = rs; S.rows = rows;
  S.random_read = random_read; S.random_write = random_write; S.first_run = first_run; S.do_draw = do_draw;
  S.ss = ss; S.bins = bins; S.threads = threads;
  pthread_barrier_init(&S.bar, 0, (unsigned)threads);
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  ref_mt_job* jobs = (ref_mt_job*)calloc((size_t)threads, sizeof(ref_mt_job));
  for (int t = 0; t < threads; ++t) {
    jobs[t].sh = &S; jobs[t].t = t; jobs[t].drawn = 0;
    pthread_create(&th[t], 0, ref_mt_worker, &jobs[t]);
  }
  uint64_t n = 0;
  for (int t = 0; t < threads; ++t) { pthread_join(th[t], 0); n += jobs[t].drawn; }
  pthread_barrier_destroy(&S.bar);
  free(th);
  free(jobs);
  return n;
}

void orc_fast_run(const orc_flame* flames, int stride, uint32_t ppt, uint64_t P, int waves, const uint16_t* perms,
                  uint64_t seed, uint32_t launch, uint32_t n_iter, int first_run, int do_draw, uint32_t W, uint32_t H,
                  const float ss[6], float* pos, double* bins, uint32_t* hits, uint64_t* drawn) {
  const uint32_t N = 64u * (uint32_t)waves;
  float* px = (float*)malloc(sizeof(float) * N * 3);
  float* nx = (float*)malloc(sizeof(float) * N * 3);
  orc_jsf* lr = (orc_jsf*)malloc(sizeof(orc_jsf) * N);
  orc_jsf wr[16];
  uint64_t cnt = 0;
  orc_flame* prep = (orc_flame*)malloc(sizeof(orc_flame));
  int prepared_for = -1;
  uint32_t mx = ppt;
  if (ppt % 2 != 0) { uint32_t v = ppt - 1; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; mx = v + 1; }
  float inv = 1.0f / (float)mx;
  uint32_t rb = 0, v = mx;
  while (v >>= 1) ++rb;
  for (uint64_t wg = 0; wg < P / N; ++wg) {
    const uint32_t base = (uint32_t)(wg * N);
    const int fi = stride ? (int)(base / ppt) : 0;
    if (fi != prepared_for) { prepare_flame(flames + fi, prep); prepared_for = fi; }
    const orc_flame* f = prep;
    orc_jsf gr;
    orc_jsf_from_philox(&gr, seed, (uint32_t)wg, launch, 2u);
    for (int w = 0; w < waves; ++w) orc_jsf_from_philox(&wr[w], seed, (uint32_t)wg * (uint32_t)waves + (uint32_t)w, launch, 1u);
    for (uint32_t t = 0; t < N; ++t) {
      const uint32_t pid = base + t;
      orc_jsf_from_philox(&lr[t], seed, pid, launch, 0u);
      if (first_run) {
        uint32_t i = pid % ppt;
        uint32_t fl = rb ? rev32(i) >> (32 - rb) : 0;
        float x = ((float)i * inv) * 2.0f - 1.0f, y = ((float)fl * inv) * 2.0f - 1.0f;
        float r1 = orc_jsf_randf(&lr[t]), r2 = orc_jsf_randf(&lr[t]);
        float s = ((r1 * 0.1f) * PI_F) * 2.0f, sn, cs;
        sincos01(sqrtf(r2), &sn, &cs);
        px[3 * t] = x + s * sn; px[3 * t + 1] = y + s * cs; px[3 * t + 2] = 0.0f;
      } else {
        px[3 * t] = pos[4 * (size_t)pid]; px[3 * t + 1] = pos[4 * (size_t)pid + 1]; px[3 * t + 2] = pos[4 * (size_t)pid + 2];
      }
    }
    for (uint32_t it = 0; it < n_iter; ++it) {
      for (int w = 0; w < waves; ++w) {
        const int xid = orc_select(f, orc_jsf_randf(&wr[w]));
        for (uint32_t l = 0; l < 64u; ++l) {
          const uint32_t t = (uint32_t)w * 64u + l;
          orc_rng rng = {0, &lr[t], 0, 0};
          float o[4];
          apply_f32_sorted(&f->xf[xid], px[3 * t], px[3 * t + 1], px[3 * t + 2], first_run && it == 0, &rng, o);
          if (badval(o[0]) || badval(o[1])) {
            o[0] = orc_jsf_randf(&lr[t]) * 2.0f - 1.0f;
            o[1] = orc_jsf_randf(&lr[t]) * 2.0f - 1.0f;
            o[3] = 0.0f;
          }
          px[3 * t] = o[0]; px[3 * t + 1] = o[1]; px[3 * t + 2] = o[2];
          if (do_draw) cnt += (uint64_t)draw_f32(f, o, &rng, W, H, ss, bins, hits);
        }
      }
      const uint32_t row = orc_jsf_next_a(&gr) & (ORC_NPERM - 1u);
      for (uint32_t t = 0; t < N; ++t) {
        const uint32_t d = perms[(size_t)row * N + t];
        nx[3 * d] = px[3 * t]; nx[3 * d + 1] = px[3 * t + 1]; nx[3 * d + 2] = px[3 * t + 2];
      }
      float* tmp = px; px = nx; nx = tmp;
    }
    for (uint32_t t = 0; t < N; ++t) {
      const uint32_t pid = base + t;
      pos[4 * (size_t)pid] = px[3 * t]; pos[4 * (size_t)pid + 1] = px[3 * t + 1];
      pos[4 * (size_t)pid + 2] = px[3 * t + 2]; pos[4 * (size_t)pid + 3] = 0.0f;
    }
  }
  free(px); free(nx); free(lr); free(prep);
  if (drawn) *drawn = cnt;
}

/* ======================================================== f64 chaos oracle */
typedef struct {
  const orc_flame* f;
  double cdf[ORC_MAX_XF];
  uint32_t W, H;
  const double* ss;
  uint64_t p0, p1;
  uint32_t fuse, iters;
  uint64_t seed;
  double* bins;
  uint64_t drawn;
} chaos_job;

static void* chaos_worker(void* arg) {
  chaos_job* J = (chaos_job*)arg;
  const orc_flame* f = J->f;
  xo_state xs;
  uint64_t sd = J->seed ^ (J->p0 * 0x9E3779B97F4A7C15ull);
  for (int i = 0; i < 4; ++i) xs.s[i] = sm64(&sd);
  orc_rng rng = {2, (orc_jsf*)&xs, 0, 0};
  uint64_t drawn = 0;
  for (uint64_t p = J->p0; p < J->p1; ++p) {
    double x = xo_unit(xs.s) * 2.0 - 1.0, y = xo_unit(xs.s) * 2.0 - 1.0, c = xo_unit(xs.s);
    for (uint32_t it = 0; it < J->fuse + J->iters; ++it) {
      double u = xo_unit(xs.s);
      int k = 0;
      while (k < f->n_xforms - 1 && u > J->cdf[k]) ++k;
      double o[4];
      apply_f64_sorted(&f->xf[k], x, y, c, 0, &rng, o);
      if (badval(o[0]) || badval(o[1])) {
        x = xo_unit(xs.s) * 2.0 - 1.0; y = xo_unit(xs.s) * 2.0 - 1.0; c = o[2];
        continue;
      }
      x = o[0]; y = o[1]; c = o[2];
      if (it < J->fuse) continue;
      double rx = o[0], ry = o[1], rc = o[2], rw = o[3];
      if (f->has_final) {
        double q[4];
        apply_f64_sorted(&f->fin, o[0], o[1], o[2], 0, &rng, q);
        rx = q[0]; ry = q[1]; rc = q[2]; rw = q[3] * o[3];
      }
      const double* s = J->ss;
      double sx = s[0] * rx + s[2] * ry + s[4], sy = s[1] * rx + s[3] * ry + s[5];
      double fx = floor(sx), fy = floor(sy);
      if (!(fx >= 0 && fy >= 0 && fx < J->W && fy < J->H && rw > 0)) continue;
      size_t b = (size_t)(J->H - (uint32_t)fy - 1u) * J->W + (uint32_t)fx;
      double t = ceil(rc * 255.0);
      int pi = !(t > 0) ? 0 : (t >= 255 ? 255 : (int)t);
      J->bins[4 * b + 0] += f->pal[pi][0];
      J->bins[4 * b + 1] += f->pal[pi][1];
      J->bins[4 * b + 2] += f->pal[pi][2];
      J->bins[4 * b + 3] += rw;
      ++drawn;
    }
  }
  J->drawn = drawn;
  return 0;
}

uint64_t orc_chaos_f64(const orc_flame* f_in, uint32_t W, uint32_t H, const double ss[6], uint64_t particles,
                       uint32_t fuse, uint32_t iters, uint64_t seed, int threads, double* bins) {
  orc_flame* f = (orc_flame*)malloc(sizeof(orc_flame));
  prepare_flame(f_in, f);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  chaos_job* jobs = (chaos_job*)calloc((size_t)threads, sizeof(chaos_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  double total = 0;
  for (int i = 0; i < f->n_xforms; ++i) total += f->xf[i].weight;
  double acc = 0;
  double cdf[ORC_MAX_XF];
  for (int i = 0; i < f->n_xforms; ++i) { acc += f->xf[i].weight / total; cdf[i] = acc; }
  const size_t nb = (size_t)W * H * 4;
  for (int t = 0; t < threads; ++t) {
    chaos_job* J = &jobs[t];
    J->f = f;
    memcpy(J->cdf, cdf, sizeof(cdf));
    J->W = W; J->H = H; J->ss = ss;
    J->p0 = particles * (uint64_t)t / (uint64_t)threads;
    J->p1 = particles * (uint64_t)(t + 1) / (uint64_t)threads;
    J->fuse = fuse; J->iters = iters; J->seed = seed;
    J->bins = (t == 0) ? bins : (double*)calloc(nb, sizeof(double));
    pthread_create(&th[t], 0, chaos_worker, J);
  }
  uint64_t drawn = 0;
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], 0);
    drawn += jobs[t].drawn;
    if (t > 0) {
      for (size_t i = 0; i < nb; ++i) bins[i] += jobs[t].bins[i];
      free(jobs[t].bins);
    }
  }
  free(jobs);
  free(th);
  free(f);
  return drawn;
}

/* ========================================================= DE and tonemap */
void orc_density(const float* hist, uint32_t W, uint32_t H, int R, int rmin, float curve, int fp32, double* out) {
  for (uint32_t py = 0; py < H; ++py) {
    for (uint32_t px = 1; px < W; ++px) {   /* px = 0: centre outside the clip volume */
      const float* b = hist + 4 * ((size_t)py * W + px);
      if (b[3] == 0.0f) continue;            /* density_vert.glsl:30-33 */
      int r;
      if (fp32) {
        float q = (float)R / powf(b[3], curve);
        int ri = (q >= (float)R) ? R : (int)q;
        r = ri > rmin ? ri : rmin;
      } else {
        double q = (double)R / pow((double)b[3], (double)curve);
        int ri = (q >= (double)R) ? R : (int)q;
        r = ri > rmin ? ri : rmin;
      }
      if (r < 0) r = 0;
      const int cx = (int)px - 1, cy = (int)H - 1 - (int)py;
      if (r == 0) {                          /* density_frag.glsl:11-14 */
        double* o = out + 4 * ((size_t)cy * W + cx);
        for (int k = 0; k < 4; ++k) o[k] += b[k];
        continue;
      }
      const int S = 2 * r + 1;
      for (int i = -r; i <= r; ++i) {
        const int ty = cy + i;
        if (ty < 0 || ty >= (int)H) continue;
        for (int j = -r; j <= r; ++j) {
          const int tx = cx + j;
          if (tx < 0 || tx >= (int)W) continue;
          double wgt;
          if (fp32) {
            float e = 1.0f / (float)S;
            float nxf = (2.0f * (float)j + e) * e, nyf = (2.0f * (float)i + e) * e;
            float d = nxf * nxf + nyf * nyf;
            if (d > 1.0f) continue;
            wgt = (double)((1.0f - d) * (0.63661977236f / (float)(r * r)));
          } else {
            double nxd = (2.0 * j + 1.0 / S) / S, nyd = (2.0 * i + 1.0 / S) / S;
            double d = nxd * nxd + nyd * nyd;
            if (d > 1.0) continue;
            wgt = (1.0 - d) * (0.63661977236 / ((double)r * r));
          }
          double* o = out + 4 * ((size_t)ty * W + tx);
          for (int k = 0; k < 4; ++k) o[k] += b[k] * wgt;
        }
      }
    }
  }
}

void orc_tonemap(const double* in, double* out, uint64_t n, double gamma, double brightness, double vib, double k) {
  for (uint64_t i = 0; i < n; ++i) {       /* tonemap.glsl:18-39 */
    const double* c = in + 4 * i;
    double* o = out + 4 * i;
    if (!(c[3] > 0.0)) { o[0] = o[1] = o[2] = 0.0; o[3] = 1.0; continue; }
    double ls = 0.5 * brightness * log10(1.0 + c[3] * k) / c[3];
    double a = c[3] * ls, ig = 1.0 / gamma, gf = pow(a, ig) / a;
    for (int q = 0; q < 3; ++q) {
      double v = c[q] * ls;
      double m = pow(v, ig) * (1.0 - vib) + gf * v * vib;
      o[q] = m < 0 ? 0 : (m > 1 ? 1 : m);
    }
    o[3] = 1.0;
  }
}
