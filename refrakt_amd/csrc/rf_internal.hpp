// Internal interface between the host library (rf_host.cpp) and the kernel
// launchers (rf_k_compat.hip, rf_k_fast.hip, rf_k_misc.hip).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "rf_device.hpp"
#include "rf_sort.hpp"

namespace rfk {

struct CompatArgs {
  const float4* pos_in;
  float4* pos_out;
  const uint32_t* shuf;          // shuffle rows, row-major, ppt entries per row
  uint4* rs;                     // jsf32 state per slot
  const rfd::Block* blocks;      // per temporal sample
  uint32_t block_stride;         // 0: one block for all samples, 1: one per sample
  const float4* palette;
  float* bins;                   // W*H RGBA
  unsigned long long* counter;   // binned samples
  uint32_t ppt;                  // particles per temporal sample
  uint32_t T;                    // temporal samples
  uint32_t in_row, out_row;
  uint32_t random_read, random_write, first_run, do_draw;
  uint32_t W, H;
  float ss[6];                   // screen affine
};

struct FastArgs {
  float4* pos;                   // P particles (x, y, c, 0), read and written
  const rfd::Block* blocks;
  uint32_t block_stride;
  const float4* palette;
  float* bins;
  unsigned long long* counter;
  const uint16_t* perms;         // NPERM rows x (64*waves) slot permutations
  uint64_t seed;
  uint32_t launch;               // launch counter (RNG streams)
  uint32_t ppt;
  uint32_t n_iter;
  uint32_t n_batches;            // batches of 64*waves particles (set by launch_fast)
  uint32_t first_run, do_draw;
  uint32_t W, H;
  float ss[6];
  SortArgs sort;                 // RF_HIST_SORT only
};

struct DEArgs {
  const float4* hist;            // W*H bins (rows bottom-up)
  float4* de_out;                // may be null
  float4* img_out;               // may be null (tonemapped)
  uint32_t W, H;
  int32_t R, rmin;
  float curve;
  int32_t do_tonemap;
  float gamma, brightness, vibrancy, k;
  // Optional completion stamp (batched rendering): the last workgroup to
  // finish writes the wall clock to *done_ts; *tiles_done counts workgroups.
  uint32_t* tiles_done;
  unsigned long long* done_ts;
  // Work buffers: every bin's splat radius (W*H bytes, -1 empty) and the
  // splat weights of radii 0..max(R, rmin) (de_weight_bytes).
  signed char* rad;
  float* wt;
};

// Bytes of the splat-weight table of radii 0..R (sum over r of (2r+1)^2 floats).
inline size_t de_weight_bytes(int R) { return (size_t)(4 * (R + 1) * (R + 1) * (R + 1) - (R + 1)) / 3 * 4; }

constexpr int NPERM = 8;          // permutation rows of the fast engine (held in LDS)
// Draw counters: COUNTER_SLOTS slots of COUNTER_STRIDE u64 (one 128-byte
// line each; slot s holds [binned samples, LDS-table re-hits]).  Memory-side
// atomics on one address serialise (about 80 M/s): one counter for every
// wave of a 2^24-particle draw cost 1.6 ms, spread over 64 lines they cost
// nothing.  The host sums the slots.
constexpr int COUNTER_SLOTS = 64;
constexpr int COUNTER_STRIDE = 16;

#ifndef __HIPCC_RTC__
hipError_t launch_init_rs(uint4* rs, uint64_t n, hipStream_t s);
hipError_t launch_compat_pass(const CompatArgs& a, int hist_mode, hipStream_t s);
hipError_t launch_fast(const FastArgs& a, uint64_t n_particles, int waves, int hist_mode, hipStream_t s);
// Fills n u64 with zero (a kernel, not a memset: safe under graph capture).
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t s);
hipError_t launch_animate(rfd::Block* blocks, uint32_t T, float tss_width, const float* rot, hipStream_t s);
// Density estimation (+ tonemap): a.rad (W*H bytes) and a.wt
// (de_weight_bytes(max(a.R, a.rmin))) must be device buffers the launch may
// overwrite; R and rmin in [0, 100].
hipError_t launch_de_tonemap(const DEArgs& a, hipStream_t s);
hipError_t launch_tonemap(const float4* in, float4* out, uint32_t n, float gamma, float brightness,
                          float vibrancy, float k, hipStream_t s);
hipError_t launch_rgba8(const float4* in, uchar4* out, uint32_t n, hipStream_t s);
hipError_t launch_eval_xform(const rfd::Block* blk, const float* pts, const float* rands, float* out,
                             uint32_t n, hipStream_t s);
hipError_t launch_atomic_probe(float* bins, uint64_t nbins, uint64_t samples, int hist_mode,
                               const float4* palette, hipStream_t s);
#endif

}  // namespace rfk
