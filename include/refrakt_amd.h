/*
 * refrakt_amd — fractal-flame chaos game, density estimation and tonemap on
 * AMD Instinct MI355X (gfx950), behind a plain C ABI.
 *
 * The entry points replace the reference renderer's hot path
 * (untrioctium/refrakt):
 *   flame::set_sim_parameters   src/flame.hpp:77,  src/flame.cpp:105-158
 *   flame::load_flame           src/flame.hpp:84,  src/flame.cpp:160-226
 *   flame::warmup               src/flame.hpp:88,  src/flame.cpp:228-281
 *   flame::draw_to_bins         src/flame.hpp:89,  src/flame.cpp:283-330
 *   density estimation pass     src/main.cpp:490-515 (+ shaders/density_*.glsl)
 *   tonemap pass                src/main.cpp:522-535 (+ shaders/tonemap.glsl)
 *   flame_compiler registry     src/variation_table.hpp:20-59
 *
 * Conventions
 *   - Every call returns RF_OK (0) or a negative RF_ERR_* code; the message of
 *     the last failure on a context is returned by rf_last_error().  No C++
 *     exception crosses this boundary.
 *   - Device memory is owned by the library.  Host arrays passed in are owned
 *     by the caller and only read or written during the call.
 *   - Calls on one rf_ctx are serialised by the caller; distinct contexts are
 *     independent.  Each context owns one HIP stream, created blocking so it
 *     orders with the legacy default stream (where torch and hipMemcpy run).
 *   - Histograms are W*H RGBA float32 bins, row-major, rows stored bottom-up
 *     exactly like the reference's `bins` SSBO (shaders/flame.glsl:84):
 *     screen pixel (x, y) lives at bin (H-1-y)*W + x.
 *   - Images produced by density estimation / tonemap are W*H RGBA float32 in
 *     the reference's render-target layout (the texture rows of
 *     render_targets[0..1], src/main.cpp:490-535).
 */
#ifndef REFRAKT_AMD_H
#define REFRAKT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
#define RF_OK 0
#define RF_ERR_ARG -1       /* invalid argument */
#define RF_ERR_PARSE -2     /* flam3 parse error, unknown attribute or variation */
#define RF_ERR_HIP -3       /* HIP runtime error */
#define RF_ERR_STATE -4     /* call not valid in the current state */
#define RF_ERR_LIMIT -5     /* size beyond a compiled-in limit */
#define RF_ERR_IO -6        /* file could not be read */

/* --------------------------------------------------------------- limits */
#define RF_MAX_XFORMS 32          /* non-final xforms per flame */
#define RF_MAX_VARS_PER_XFORM 16  /* variation terms per xform */
#define RF_MAX_VAR_PARAMS 8       /* parameters per variation */
#define RF_NOALLOC_SCRATCH_BYTES 8192  /* device scratch of rf_draw_to_bins_noalloc, per GPU */
#define RF_PALETTE_SIZE 256

/* ---------------------------------------------------------------- engines */
/* RF_ENGINE_COMPAT reproduces the reference's pass structure exactly: one
 * launch per pass, 256-lane workgroups sharing one xform draw, per-slot
 * jsf32 state, shuffle-row gathers (shaders/flame.glsl:41-90).  With the same
 * shuffle rows and seeds it gives the reference's bin indices bit for bit on
 * affine/linear-only flames.
 * RF_ENGINE_FAST is the MI355X engine: one persistent launch runs all passes,
 * particles stay in registers, each wave draws one xform per iteration and
 * particles are exchanged across waves through LDS every iteration. */
#define RF_ENGINE_COMPAT 0
#define RF_ENGINE_FAST 1

/* Histogram accumulation strategies of the fast engine (the compat engine
 * uses QUAD for AUTO and LDS). */
#define RF_HIST_AUTO 0     /* LDS for one pass of the first draw; then SORT if under 90 % of its samples
                              re-hit the LDS tables (QUAD if SORT is not allowed and under 2 %) */
#define RF_HIST_QUAD 1     /* 4 lanes per RGBA bin: one 16-byte atomic request per sample */
#define RF_HIST_WAVEAGG 2  /* merge same-bin hits inside a wave before the atomics */
#define RF_HIST_NAIVE 3    /* 4 separate float atomics per sample (baseline) */
#define RF_HIST_LDS 4      /* per-workgroup LDS table of hot bins, flushed once per launch; misses go QUAD */
#define RF_HIST_SORT 5     /* records sorted into LDS-sized tiles, summed in LDS: no per-sample atomics;
                              needs unit opacities (else QUAD) */

typedef struct rf_ctx rf_ctx;
typedef struct rf_flame rf_flame;
typedef struct rf_hist rf_hist;

/* ------------------------------------------------------------ flame model */
/* One variation term of an xform: the reference keeps these in
 * std::map<std::string,float> (src/flame.hpp:26-27); rf_flame_create sorts
 * them by name so they are summed in the same (alphabetical) order. */
typedef struct rf_var_term {
  int32_t id;                        /* rf_variation_id(name) */
  float weight;
  float params[RF_MAX_VAR_PARAMS];   /* in rf_variation_param_name order */
} rf_var_term;

/* flame_xform (src/flame.hpp:21-37). affine/post are flam3 `coefs` order
 * a b c d e f:  x' = a*x + c*y + e,  y' = b*x + d*y + f. */
typedef struct rf_xform {
  float affine[6];
  float post[6];
  int32_t has_post;
  int32_t n_vars;
  rf_var_term vars[RF_MAX_VARS_PER_XFORM];
  float weight;
  float color;
  float color_speed;
  float rotation_frequency;
  float opacity;
} rf_xform;

/* flame (src/flame.hpp:41-60). */
typedef struct rf_flame_desc {
  int32_t n_xforms;
  rf_xform xforms[RF_MAX_XFORMS];
  int32_t has_final;
  rf_xform final_xform;
  float palette[RF_PALETTE_SIZE][4];
  uint32_t size[2];
  float center[2];
  float scale;
  float rotate;
  int32_t estimator_min;
  int32_t estimator_radius;
  float estimator_curve;
  float gamma;
  float vibrancy;
  float brightness;
} rf_flame_desc;

/* set_sim_parameters (src/flame.cpp:105-158) plus this library's choices. */
typedef struct rf_sim_params {
  uint64_t total_particles;   /* P; a multiple of 256 * temporal_samples, below 2^32 */
  uint32_t temporal_samples;  /* T >= 1.  The reference animates T/32 groups of 32 samples
                                 (shaders/templates/animate.tpl.glsl); here any T works and
                                 T = 1 is the centre sample (no rotation). */
  uint32_t shuffle_count;     /* shuffle rows (compat engine); 0 -> 1024 */
  uint64_t seed;              /* seeds RNG states, shuffle rows and the row schedule */
  int32_t device;             /* HIP device ordinal */
  int32_t engine;             /* RF_ENGINE_* */
  int32_t hist_mode;          /* RF_HIST_* (fast engine) */
  int32_t block_waves;        /* fast engine: waves per workgroup (4 or 8); 0 -> 8 */
} rf_sim_params;

/* Density estimation parameters (src/main.cpp:500-511). */
typedef struct rf_de_params {
  int32_t estimator_radius;   /* clamped to 100 like src/main.cpp:502 */
  int32_t estimator_min;
  float estimator_curve;
} rf_de_params;

/* Tonemap parameters (shaders/tonemap.glsl:9-12, src/main.cpp:522-531). */
typedef struct rf_tonemap_params {
  float gamma;
  float brightness;
  float vibrancy;
  float scale_constant;       /* k in log10(1 + a*k); the reference passes 1/10^4 */
} rf_tonemap_params;

typedef struct rf_render_params {
  uint32_t width, height;     /* accumulator size (image size x supersample) */
  uint32_t warmup_passes;     /* passes after the first run, not binned */
  uint32_t draw_passes;       /* binned passes per draw call */
  uint32_t draw_calls;        /* number of rf_draw_to_bins calls */
  float tss_width;            /* temporal sample width (src/main.cpp:236) */
  int32_t density_estimation; /* 0: skip DE (tonemap the histogram directly) */
  rf_de_params de;            /* ignored unless density_estimation */
  rf_tonemap_params tm;
} rf_render_params;

typedef struct rf_stats {
  uint64_t drawn;             /* binned samples (the reference's atomic counter) */
  uint64_t iterations;        /* particle iterations, binned or not */
  double ms_warmup;           /* host clock: rf_warmup */
  double ms_draw;             /* host clock: the draw_calls rf_draw_to_bins calls */
  double ms_density;          /* device time of the density-estimation kernel */
  double ms_tonemap;          /* device time of the tonemap kernel */
} rf_stats;

/* ------------------------------------------------------------- context */
/* ≙ flame::set_sim_parameters + GL context creation. */
int rf_create(const rf_sim_params* params, rf_ctx** out);
void rf_destroy(rf_ctx* ctx);
/* ≙ flame::set_sim_parameters (src/flame.cpp:105-158): reallocates the
 * simulation buffers; every flame of this context needs a new warmup. */
int rf_set_sim_parameters(rf_ctx* ctx, uint64_t total_particles,
                          uint32_t temporal_samples, uint32_t shuffle_count);
int rf_set_engine(rf_ctx* ctx, int32_t engine, int32_t hist_mode);
const char* rf_last_error(const rf_ctx* ctx);  /* NULL: why the last rf_create failed */
/* Device time (ms, HIP events on the context stream) of the chaos-game
 * launches of the last rf_draw_to_bins call. */
double rf_last_draw_ms(const rf_ctx* ctx);
/* HIP stream of the context (a hipStream_t), for callers that enqueue work
 * of their own on the same device memory. */
void* rf_stream(rf_ctx* ctx);

/* ------------------------------------------------------------- flames */
/* ≙ flame::load_flame (src/flame.cpp:160-226): reads a flam3 XML file. */
int rf_flame_load(rf_ctx* ctx, const char* path, rf_flame** out);
/* Same, from XML text in memory. */
int rf_flame_parse(rf_ctx* ctx, const char* xml, size_t len, rf_flame** out);
/* Parses flam3 XML into a description without creating a flame. */
int rf_flame_parse_desc(const char* xml, size_t len, rf_flame_desc* out,
                        char* err, size_t err_len);
int rf_flame_create(rf_ctx* ctx, const rf_flame_desc* desc, rf_flame** out);
/* Replaces the flame's description (e.g. after rotating affines for
 * animation, src/main.cpp:383-395); the next draw warms up again. */
int rf_flame_update(rf_flame* flame, const rf_flame_desc* desc);
int rf_flame_get_desc(const rf_flame* flame, rf_flame_desc* out);
void rf_flame_destroy(rf_flame* flame);
/* The reference's flat parameter block (make_shader_buffer_map +
 * copy_flame_data_to_buffer, src/flame.cpp:33-103): writes at most `cap`
 * floats and returns the block size in floats (negative on error). */
int rf_flame_param_block(const rf_flame_desc* desc, float* out, int cap);
/* Checks a description the way rf_flame_create does (no GPU needed). */
int rf_flame_validate(const rf_flame_desc* desc, char* err, size_t err_len);

/* ---------------------------------------------------------- histograms */
int rf_hist_create(rf_ctx* ctx, uint32_t width, uint32_t height, rf_hist** out);
/* Wraps caller-owned device memory of width*height*16 bytes. */
int rf_hist_wrap(rf_ctx* ctx, void* device_ptr, uint32_t width, uint32_t height, rf_hist** out);
void rf_hist_destroy(rf_hist* hist);
int rf_hist_zero(rf_hist* hist);                    /* ≙ bins.zero_out() */
int rf_hist_download(rf_hist* hist, float* host);   /* W*H*4 floats */
int rf_hist_upload(rf_hist* hist, const float* host);
void* rf_hist_device_ptr(rf_hist* hist);

/* ------------------------------------------------------------ hot path */
/* ≙ flame::warmup (src/flame.cpp:228-281): animate, one first-run pass
 * seeded from the Hammersley points, then num_passes passes; nothing binned. */
int rf_warmup(rf_ctx* ctx, rf_flame* flame, uint32_t num_passes, float tss_width);
/* ≙ flame::draw_to_bins (src/flame.cpp:283-330): num_iter binned passes;
 * *drawn receives the number of binned samples (may be NULL).  With the fast
 * engine's sort histogram a call of more than 64 passes (or 2^28 samples)
 * runs as several launches of at most that many passes, each with its own
 * RNG launch key: the same result as drawing them in separate calls. */
int rf_draw_to_bins(rf_ctx* ctx, rf_flame* flame, rf_hist* bins,
                    uint32_t num_iter, uint64_t* drawn);
/* Compat-engine variants with an explicit shuffle-row schedule: rows holds
 * (in_row, out_row) pairs, one per dispatched pass (warmup: num_passes + 1
 * pairs, the first for the first run).  The reference draws these with an
 * unseeded std::mt19937 (src/flame.cpp:231-233, 285-287). */
int rf_warmup_rows(rf_ctx* ctx, rf_flame* flame, uint32_t num_passes,
                   float tss_width, const uint32_t* rows);
int rf_draw_to_bins_rows(rf_ctx* ctx, rf_flame* flame, rf_hist* bins,
                         uint32_t num_iter, const uint32_t* rows, uint64_t* drawn);
/* Allocation-free rf_draw_to_bins.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Calls no
 * hipMalloc, hipFree or other allocating routine: the draw counters live in
 * a static device array of RF_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends, its address taken per device by rf_create) and are cleared
 * by a kernel on every call.  The caller passes a warmed-up flame and a
 * histogram as for rf_draw_to_bins; the fast engine only (RF_ERR_ARG
 * otherwise), in a direct histogram mode: RF_HIST_SORT, and AUTO's sort
 * choice, run the LDS bin table instead (no record buffers), and AUTO never
 * probes.  Uses a flame program only if one is already compiled (else the
 * precompiled interpreter, same results).  Results equal rf_draw_to_bins in
 * the same mode; rf_last_draw_ms reads -1 after it.  Calls that overlap on
 * one GPU (any contexts) race on the shared counters. */
int rf_draw_to_bins_noalloc(rf_ctx* ctx, rf_flame* flame, rf_hist* bins,
                            uint32_t num_iter, uint64_t* drawn);
/* Density estimation into `out` (device memory of W*H*16 bytes, or NULL to
 * use the histogram's own image buffer), ≙ src/main.cpp:490-515. */
int rf_density_estimate(rf_ctx* ctx, rf_hist* bins, const rf_de_params* de, void* out_device);
/* Tonemap `in_device` (W*H RGBA) into `out_device`, ≙ src/main.cpp:522-535. */
int rf_tonemap(rf_ctx* ctx, const void* in_device, void* out_device,
               uint32_t width, uint32_t height, const rf_tonemap_params* tm);
/* Fused density estimation + tonemap of a histogram into out_device. */
int rf_density_tonemap(rf_ctx* ctx, rf_hist* bins, const rf_de_params* de,
                       const rf_tonemap_params* tm, void* out_device);
/* The whole path (src/main.cpp:403-535): zero the context's histogram,
 * warmup, draw_calls x draw_to_bins, DE, tonemap (two kernels, as the
 * reference's two passes; the image equals rf_density_tonemap of the same
 * histogram).  hist_out and image_out are host arrays of W*H*4 floats (either
 * may be NULL). */
int rf_render(rf_ctx* ctx, rf_flame* flame, const rf_render_params* params,
              float* hist_out, float* image_out, rf_stats* stats);

/* ---------------------------------------------------------- batching */
/* Many frames (each its own flame, size and sample count) rendered in one
 * HIP graph replayed by rf_batch_run: the work items (fresh particles of one
 * frame) run in frame order in slices of at most 2^29 samples, each one
 * persistent chaos-game launch (with RF_HIST_SORT followed by its record
 * passes); the DE + tonemap of the frames a slice completes runs on a second
 * stream while the next slice draws.  This
 * generalises the reference's temporal-sample axis (one dispatch over T
 * parameter blocks, src/flame.cpp:275, shaders/templates/animate.tpl.glsl)
 * to frames of uneven size and xform count.  No reference counterpart. */
typedef struct rf_batch rf_batch;

typedef struct rf_frame_desc {
  const rf_flame_desc* flame;  /* read during rf_batch_create only */
  uint32_t width, height;      /* accumulator and image size */
  uint64_t samples;            /* binned samples wanted, e.g. spp * width * height */
} rf_frame_desc;

typedef struct rf_batch_config {
  int32_t block_waves;         /* 4 or 8: a work item is 64 * block_waves particles */
  int32_t hist_mode;           /* RF_HIST_QUAD, RF_HIST_LDS or RF_HIST_SORT (sort needs unit
                                  opacities and palettes in [0, 1], else LDS or quad run;
                                  its record capacities come from the last finished replay) */
  uint32_t iters;              /* binned iterations per particle */
  uint32_t fuse;               /* unbinned iterations per particle before those */
} rf_batch_config;

int rf_batch_create(rf_ctx* ctx, const rf_frame_desc* frames, uint32_t n_frames, rf_batch** out);
void rf_batch_destroy(rf_batch* batch);
/* Sets the launch shape and captures the graph (default: 8 waves, LDS, 256, 20). */
int rf_batch_set_config(rf_batch* batch, const rf_batch_config* cfg);
int rf_batch_get_config(const rf_batch* batch, rf_batch_config* cfg);
/* Autotuner: times every candidate shape (block_waves x hist_mode x iters) on
 * a copy of the batch whose sample counts are scaled by sample_fraction
 * (0 < f <= 1), then sets the fastest.  best/best_ms may be NULL. */
int rf_batch_tune(rf_batch* batch, double sample_fraction, rf_batch_config* best, double* best_ms);
/* 1 if the batch's chaos-game launch runs a flame program (every frame's
 * xform i has the structure of the same xform of the frame with the most
 * xforms, and flame programs are on), 0 if it is interpreted;
 * *compile_seconds (may be NULL) = seconds spent compiling it. */
int rf_batch_program_info(const rf_batch* batch, double* compile_seconds);
/* Replays the graph: zero, chaos game, DE + tonemap of every frame; *ms is
 * the host wall time of the replay. */
int rf_batch_run(rf_batch* batch, double* ms);
/* Per-frame results of the last run (arrays of n_frames, NULL to skip):
 * binned samples, and the device-clock times, in ms from the start of the
 * replay, at which the frame's last samples reached its histogram and its
 * tonemapped image was complete. */
int rf_batch_frame_stats(const rf_batch* batch, uint64_t* drawn, double* chaos_done_ms, double* image_done_ms);
/* Device pointers of frame i's histogram (bins layout) and tonemapped image. */
void* rf_batch_hist_ptr(rf_batch* batch, uint32_t i);
void* rf_batch_image_ptr(rf_batch* batch, uint32_t i);
/* Copies frame i's histogram (which 0) or image (which 1) to host (W*H*4 floats). */
int rf_batch_download(rf_batch* batch, uint32_t i, int32_t which, float* host);

/* ---------------------------------------------------------- denoiser */
/* A learned stand-in for density estimation (src/main.cpp:490-515): a
 * 4-level U-Net (channels 32-256, 1.18 M parameters, bf16 activations with
 * fp32 accumulation on MFMA, fp32 master weights, Adam) that maps the
 * normalised log-density image of a low-sample histogram to that of a
 * high-sample one.  Tensors are device memory, NHWC:
 *   x, target: N x H x W x 16 bf16 (channels 0..3 = log10(1 + v * W*H / n)
 *              of the RGBA histogram, bins layout, channel 4 the noise level
 *              log10(1 + W*H / n), from rf_log_density_input; channels
 *              5..15 zero), H and W multiples of 8;
 *   y:         N x H x W x 4 float (the denoised log-density).
 * No reference counterpart. */
typedef struct rf_denoiser rf_denoiser;
int rf_denoiser_create(rf_ctx* ctx, uint64_t seed, rf_denoiser** out);
void rf_denoiser_destroy(rf_denoiser* dn);
size_t rf_denoiser_num_params(const rf_denoiser* dn);
int rf_denoiser_get_params(rf_denoiser* dn, float* host, size_t n);
int rf_denoiser_set_params(rf_denoiser* dn, const float* host, size_t n);   /* also resets Adam */
/* One Adam step on the mean squared error over channels 0..3; *loss (may be
 * NULL, which avoids a host sync) receives the batch MSE before the step. */
int rf_denoiser_train_step(rf_denoiser* dn, const void* x, const void* target, int32_t N, int32_t H, int32_t W,
                           float lr, float* loss);
int rf_denoiser_infer(rf_denoiser* dn, const void* x, float* y, int32_t N, int32_t H, int32_t W);
/* Gradients of the MSE w.r.t. the flat fp32 parameters (host array of
 * rf_denoiser_num_params floats), without updating them. */
int rf_denoiser_grads(rf_denoiser* dn, const void* x, const void* target, int32_t N, int32_t H, int32_t W,
                      float* grads, float* loss);
/* x input of the denoiser from a device histogram (npix RGBA floats):
 * out[p, c] = log10(1 + hist[p, c] * scale) for c < 4,
 * out[p, 4] = log10(1 + scale), 0 for c in 5..15. */
int rf_log_density_input(rf_ctx* ctx, const void* hist, void* out, uint32_t npix, double scale);
/* Kernel-level entry points of the network (tests).  Views: channel stride
 * cs and first channel co of an NHWC bf16 tensor; weights bf16
 * [3][3][Cout][Cin]; dw fp32 [3][3][Cout][Cin] and db fp32 [Cout] accumulate. */
int rf_nn_conv3x3(rf_ctx* ctx, const void* in, int32_t in_cs, int32_t in_co, const void* w, const float* bias,
                  void* out, int32_t out_cs, int32_t out_co, int32_t N, int32_t H, int32_t W, int32_t Cin,
                  int32_t Cout, int32_t relu);
int rf_nn_conv3x3_wgrad(rf_ctx* ctx, const void* x, int32_t x_cs, int32_t x_co, const void* dy, int32_t dy_cs,
                        int32_t dy_co, float* dw, float* db, int32_t N, int32_t H, int32_t W, int32_t Cin,
                        int32_t Cout);
int rf_nn_flip_weights(rf_ctx* ctx, const void* w, void* wt, int32_t Cin, int32_t Cout);
int rf_nn_adam(rf_ctx* ctx, float* p, const float* g, float* m, float* v, int32_t n, float lr, float b1, float b2,
               float eps, int32_t step);

/* ---------------------------------------------------------- output */
/* RGBA float -> RGBA8 the way the reference reads its render target back for
 * a screenshot (src/buffer_objects.hpp:117: clamp to [0,1], x255, round to
 * nearest; NaN -> 0).  Host arrays of W*H*4. */
int rf_rgba8_from_f32(const float* in, uint32_t width, uint32_t height, uint8_t* out);
/* The same conversion on the GPU: device pointers, on the context's stream. */
int rf_rgba8_device(rf_ctx* ctx, const void* in_device, void* out_device, uint32_t width, uint32_t height);
/* Writes W*H RGBA8 pixels (first row first) as a PNG file, ≙ the screenshot
 * of src/main.cpp:590-593. */
int rf_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height);
/* Writes W*H RGBA float pixels (first row first) as an uncompressed
 * single-part scanline OpenEXR file, channels stored as HALF (half != 0) or
 * FLOAT.  No counterpart in the reference, which saves 8-bit PNG only
 * (src/main.cpp:590-593): for the density-estimated or tonemapped float
 * image without quantisation. */
int rf_write_exr(const char* path, const float* rgba, uint32_t width, uint32_t height, int32_t half);
/* IEEE binary32 -> binary16, round to nearest even (the EXR HALF encoding). */
uint16_t rf_half_from_float(float f);

/* ---------------------------------------------------- state inspection */
/* Particle buffer (P x (x, y, color, 0)) after the last pass. */
int rf_particles_download(rf_ctx* ctx, rf_flame* flame, float* host);
/* Compat engine: jsf32 state per slot (P x 4 u32), shuffle rows
 * (shuffle_count x P/T u32) and Hammersley points (P/T x 4 floats). */
int rf_rng_download(rf_ctx* ctx, uint32_t* host);
int rf_shuffle_download(rf_ctx* ctx, uint32_t* host);
int rf_sample_points_download(rf_ctx* ctx, float* host);
/* Fast engine: the 8 LDS permutation rows (8 x 64*block_waves u16). */
int rf_fast_perms_download(rf_ctx* ctx, uint16_t* host);
/* Histogram strategy the flame's next fast-engine draw uses (resolves
 * RF_HIST_AUTO) and, if a draw has used RF_HIST_LDS, the fraction of its binned
 * samples whose bin was already in their workgroup's LDS table (else -1). */
int rf_flame_hist_mode(const rf_flame* flame, double* lds_hit_fraction);
/* Fraction of the binned samples of the flame's last RF_HIST_SORT draw that
 * overflowed the bucket/tile capacities (and were added with atomics), or -1. */
double rf_flame_sort_overflow(const rf_flame* flame);
/* --------------------------------------------------- flame programs */
/* The reference generates and compiles a GLSL program per flame
 * (src/variation_table.cpp:217-265, compiled in src/flame.cpp:285-288).  The
 * fast engine's precompiled kernels interpret any flame; with flame programs
 * on (the default; environment RF_FLAME_PROGRAMS=0 turns them off for new
 * contexts) the first warmup or draw of a flame compiles the same kernel with
 * the flame's structure fixed (hiprtc, cached per structure and device for the
 * life of the process) and later launches run it.  Results are the same bits
 * as the interpreter's.  enable: 1 on, 0 off. */
int rf_set_flame_programs(rf_ctx* ctx, int32_t enable);
/* 1 if the flame's last fast-engine launch ran a flame program, 0 if it was
 * interpreted, -1 if none ran yet; *compile_seconds (may be NULL) = seconds
 * spent compiling its program in this process (0 when it came from the cache). */
int rf_flame_program_info(const rf_flame* flame, double* compile_seconds);
/* HIP source of the flame program of desc with one entry point for
 * (block_waves, hist_mode): hist_mode 1..5 for the fast engine, 16 + 1,
 * 16 + 4 or 16 + 5 for the batched renderer's kernel.  Writes at most cap bytes
 * (NUL-terminated) and returns the full length + 1, or < 0 on error. */
int rf_flame_program_source(const rf_flame_desc* desc, int32_t block_waves, int32_t hist_mode, char* buf,
                            size_t cap);
/* Compiles that source for arch (e.g. "gfx950") without a device: the code
 * object goes to code (cap bytes; *size = its size), the compiler log to log.
 * *seconds (may be NULL) = compile time.  RF_OK, or RF_ERR_ARG / RF_ERR_HIP. */
int rf_flame_program_build(const rf_flame_desc* desc, int32_t block_waves, int32_t hist_mode, const char* arch,
                           void* code, size_t cap, size_t* size, char* log, size_t log_len, double* seconds);

/* Screen-space affine used by draw_to_bins (src/flame.cpp:289-296). */
int rf_screen_affine(const rf_flame_desc* desc, uint32_t width, uint32_t height, float out[6]);

/* ------------------------------------------------ variation registry */
/* ≙ flame_compiler (src/variation_table.hpp:20-59). */
int rf_variation_count(void);
const char* rf_variation_name(int id);
int rf_variation_id(const char* name);               /* -1 if unknown */
int rf_variation_num_params(int id);
const char* rf_variation_param_name(int id, int k);
/* Bit 0: no_weight_mul, bit 1: pre_xform, bit 8: consumes random numbers. */
int rf_variation_flags(int id);

/* ------------------------------------------------- diagnostics / tests */
/* Applies one xform (affine, variations, post, colour blend) to n points on
 * the GPU.  pts: n x (x, y, c); rands: n x 16 uniforms consumed in order by
 * RAND(); out: n x (x, y, c, opacity).  Host arrays. */
int rf_eval_xform(rf_ctx* ctx, const rf_xform* xf, const float* pts,
                  const float* rands, float* out, uint32_t n);
/* Measures the L2 float-atomic ceiling: `ops` RGBA samples scattered over a
 * footprint of `footprint_bins` bins with the given strategy; returns the
 * kernel time in ms through *ms. */
int rf_atomic_probe(rf_ctx* ctx, int32_t hist_mode, uint64_t footprint_bins,
                    uint64_t samples, double* ms);

/* sizeof of the ABI structs (0 rf_var_term, 1 rf_xform, 2 rf_flame_desc,
 * 3 rf_sim_params, 4 rf_de_params, 5 rf_tonemap_params, 6 rf_render_params,
 * 7 rf_stats, 8 rf_frame_desc, 9 rf_batch_config), for bindings that check
 * their layouts. */
size_t rf_struct_size(int which);

#ifdef __cplusplus
}
#endif
#endif /* REFRAKT_AMD_H */
