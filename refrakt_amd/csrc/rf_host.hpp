// Host-side objects behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/refrakt_amd.h"
#include "rf_device.hpp"

struct rf_ctx;
namespace rfj {
struct Program;
}

struct rf_var_info {
  const char* name;
  int n_params;
  const char* params[RF_MAX_VAR_PARAMS];
  uint32_t flags;     // 1 no_weight_mul, 2 pre_xform
  uint32_t precalc;
  int uses_rand;
  int ref_line;       // line of the entry in the reference's variations.yaml
};

namespace rfh {

int parse_flam3(const char* xml, size_t len, rf_flame_desc* out, std::string& err);
bool param_lookup(const char* name, int* var, int* k);
const rf_var_info* var_info(int id);

// Device buffer owned by the library.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t alloc(size_t n);
  void release();
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Pinned host buffer (hipHostMalloc): device-to-host copies into it run
// asynchronously, so a draw's readbacks share one stream synchronisation.
struct HostBuf {
  void* p = nullptr;
  void* dev = nullptr;   // the same memory's address for kernels
  size_t bytes = 0;
  hipError_t alloc(size_t n);
  void release();
  ~HostBuf() { release(); }
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
};

// Builds the device Block of a flame (weights normalised over the non-final
// xforms, terms in name order, pre_xform terms first).  Returns false and
// sets err on a limit violation.
bool build_block(const rf_flame_desc& d, rfd::Block& b, std::string& err);
// The reference's flat parameter block (src/flame.cpp:33-103).
std::vector<float> reference_param_block(const rf_flame_desc& d);
// Screen-space affine (src/flame.cpp:289-296, src/flame.hpp:97-128).
void screen_affine(const rf_flame_desc& d, uint32_t w, uint32_t h, float out[6]);
// The fast engine's 16 LDS permutation rows for 64*waves lanes.
std::vector<uint16_t> perm_rows(uint64_t seed, int waves);
// True when every palette channel and every opacity lies in [0, 1]: the LDS
// histogram paths accumulate in fixed point and need that range.
bool fixed_point_ok(const rf_flame_desc& d);
// Records msg as the context's last error and returns code.
int set_err(rf_ctx* ctx, int code, const std::string& msg);
// Orders the tile-summing segments (records a segment can hold, job) largest
// first, in steps of 1024 records and stable within a step, so the long
// workgroups start early: a counting pass instead of a comparison sort.
void order_segments(std::vector<std::pair<uint32_t, uint32_t>>& segs);

}  // namespace rfh

#define HIPCHK(ctx, expr)                                                                         \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      return rfh::set_err((ctx), RF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

struct rf_ctx {
  rf_sim_params sp{};
  hipStream_t stream = nullptr;
  std::string err;
  uint64_t ppt = 0;            // particles per temporal sample
  uint32_t shuffle_count = 0;
  uint64_t generation = 1;     // bumped by rf_set_sim_parameters
  uint32_t launch = 0;         // fast-engine launch counter
  uint64_t row_state = 0;      // compat-engine shuffle-row schedule stream
  // compat engine buffers
  rfh::DevBuf swap, rs, shuf, samples;
  bool compat_ready = false;
  // fast engine
  rfh::DevBuf perms;
  int perm_waves = 0;
  rfh::DevBuf counter;         // u64 [drawn, samples added in the LDS bin cache]
  // RF_HIST_SORT buffers: bucket and tile records, cursors, offsets, capacities
  rfh::DevBuf sort_cbuf, sort_tbuf, sort_bcur, sort_tcur, sort_mass;
  // a draw's readbacks: counter slots, then (sort draws) bucket and tile
  // cursors, written by one kernel
  rfh::HostBuf readback;
  // staging of a sort draw's capacity plan: an upload from pinned memory
  // does not wait for the copy the way one from pageable memory does
  rfh::HostBuf meta_stage;
  rfh::DevBuf probe;           // atomic probe scratch
  rf_hist* render_hist = nullptr;
  rfh::DevBuf render_img;
  rfh::DevBuf de_wt;           // density estimation: splat weight table
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_draw_ms = 0.0;   // device time of the last draw_to_bins' chaos-game launches
  bool programs = true;        // fast engine: run flame programs (rf_jit.cpp)
};

struct rf_flame {
  rf_ctx* ctx = nullptr;
  rf_flame_desc desc{};
  uint64_t generation = 0;     // ctx generation the particle state belongs to
  bool needs_update = true;
  bool animated = false;
  uint32_t n_blocks = 1;       // blocks read by the kernels (1 or T)
  rfh::DevBuf blocks;          // n_blocks (+1 pristine when animated) Blocks
  rfh::DevBuf rot;             // rotation frequencies (animated)
  rfh::DevBuf palette;         // 256 float4
  rfh::DevBuf local;           // particles
  bool have_particles = false;
  // RF_HIST_AUTO: the strategy the next draw uses.  It starts with the LDS
  // bin cache and falls back to the quad path for the rest of the flame's
  // life (until its next upload) if the first draw added fewer than
  // kAutoLdsMinHits of its samples in LDS.
  int auto_mode = RF_HIST_LDS;
  double last_lds_hit_fraction = -1.0;
  bool auto_probed = false;  // AUTO's first draw has measured the LDS table on one pass
  // RF_HIST_SORT: fraction of the samples that went to each bucket and tile
  // in the last draw (sizes the next draw's capacities), and the fraction of
  // binned samples that overflowed them.
  uint32_t sort_bins = 0;
  std::vector<double> sort_bfrac, sort_tfrac;
  double last_sort_overflow = -1.0;
  // The last sort plan's capacities and work lists (host copy and device
  // copy), made for sort_meta_S samples (up to twice the draw that made it)
  // and reused by draws of a quarter of that up to all of it while the
  // histogram stays the same and nothing overflows.
  std::vector<uint32_t> sort_meta;
  rfh::DevBuf sort_meta_dev;
  uint64_t sort_meta_S = 0;
  uint32_t sort_meta_bins = 0, sort_meta_chunks = 0, sort_meta_segs = 0;
  size_t sort_meta_fixed = 0;
  // Flame program: the host copy of the Block (its structure selects the
  // program), whether the last fast-engine launch ran one.
  std::vector<rfd::Block> host_block;
  // the program of host_block's structure once looked up (programs live as
  // long as the process); cleared when the flame is uploaded again
  const rfj::Program* program = nullptr;
  int program_used = -1;
  double program_compile_s = 0.0;
};

struct rf_hist {
  rf_ctx* ctx = nullptr;
  uint32_t w = 0, h = 0;
  void* ptr = nullptr;
  bool owned = false;
  rfh::DevBuf own;
  rfh::DevBuf image;           // density-estimated / tonemapped output
  rfh::DevBuf de_rad;          // density estimation: splat radius per bin (W*H bytes)
};
