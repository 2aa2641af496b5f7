// Flame programs: the fast engine's kernel compiled per flame structure.
//
// The reference generates and compiles one GLSL dispatch() per flame
// (src/variation_table.cpp:217-265, src/flame.cpp:285-288).  Here the
// precompiled kernels interpret a flame from its Block; a flame program is
// the same kernel body (rf_k_fast_body.hpp) instantiated with the flame's
// structure -- which variations each xform runs, its precalcs, post-affine and
// final xform -- fixed at compile time with hiprtc, so the per-term variation
// switch, the term loop and their scalar loads disappear.  Coefficients,
// weights and parameters stay in the Block: a flame whose values change but
// whose structure does not keeps its program.  Programs are cached per
// structure and device for the life of the process.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rf_device.hpp"
#include "rf_batch.hpp"
#include "rf_internal.hpp"

namespace rfj {

struct Program;

// Kernels of a program are (waves, mode) pairs: mode 1..5 is the fast
// engine with that histogram mode, kBatchKernel + mode the batched renderer
// (modes 1 and 4).
constexpr int kBatchKernel = 16;

// Text identifying a Block's structure (programs are shared between flames
// with equal keys).
std::string structure_key(const rfd::Block& b);

// HIP source of the program for `b` with one entry point per (waves, hist
// mode) pair in `kernels`.
std::string program_source(const rfd::Block& b, const std::vector<std::pair<int, int>>& kernels);

// Name of the entry point for (waves, hist mode).
std::string entry_name(int waves, int hist_mode);

// Compiles `src` for `arch` (e.g. "gfx950") with hiprtc; returns false and
// sets log on failure.
bool compile(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& log,
             double* seconds);

// The program for b's structure on the current device with entry points for
// every (waves, hist mode) in `kernels`, compiled and loaded on first use.
// Returns nullptr and sets err on failure.
const Program* get_program(const rfd::Block& b, const std::vector<std::pair<int, int>>& kernels,
                           std::string& err);

// The already compiled program for b's structure on the current device if it
// has the entry point `kernel`; nullptr otherwise.  Never compiles or loads.
const Program* find_program(const rfd::Block& b, std::pair<int, int> kernel);

// True when p was built for `device` and has an entry point for every
// (waves, hist mode) in `kernels`: a cached program pointer still serves.
bool program_has(const Program* p, int device, const std::vector<std::pair<int, int>>& kernels);

// Seconds spent compiling the program (0 if it came from the cache).
double program_compile_seconds(const Program* p);

// Launches the program's fast-engine kernel over n particles, like
// rfk::launch_fast for the precompiled kernels.
hipError_t launch_fast(const Program* p, const rfk::FastArgs& a, uint64_t n, int waves, int hist_mode,
                       hipStream_t s);

// Launches the program's batched-rendering kernel (one workgroup per
// resident slot, at most max_grid if non-zero), like rfk::launch_batch.
hipError_t launch_batch(const Program* p, const rfk::BatchArgs& a, int waves, int hist_mode, uint32_t max_grid,
                        hipStream_t s);

// True when every xform of `b` has the structure of the same xform of `big`
// and both have the same final xform: a program for `big` runs `b`.
bool structure_covers(const rfd::Block& big, const rfd::Block& b);

}  // namespace rfj
