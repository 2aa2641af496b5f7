// Flame programs: source generation, hiprtc compilation, module cache, launch.
#include "rf_jit.hpp"

#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

// k_jit_headers[]: {include name, text} of the device headers the program
// source includes, embedded at build time (refrakt_amd/build.py).
#include "gen/jit_headers.inc"

namespace rfj {
namespace {

// The LDS bin-cache table size per workgroup, as the precompiled kernels use.
int lds_logk(int waves, int mode) {
  if (mode >= kBatchKernel) return mode - kBatchKernel == 4 ? (waves == 8 ? 11 : 10) : 0;
  return mode == 4 ? 10 : 0;
}

std::string xform_key(const rfd::Block& b, int xi) {
  const rfd::Xform& X = b.xf[xi];
  std::ostringstream o;
  o << X.precalc << "," << (X.flags & rfd::XF_HAS_POST) << "," << X.n_pre;
  for (uint32_t k = 0; k < X.n_terms; ++k) {
    const rfd::Term& T = b.terms[X.term_begin + k];
    o << "," << (T.var | (T.flags << 16));
  }
  return o.str();
}

// An xform's terms start after those of the xforms before it: the same in
// every flame with this structure, so the offset is a constant, except for
// the final xform of the batched renderer (load_tb), whose programs also run
// flames with fewer xforms (structure_covers).
void xform_args(std::ostringstream& o, const rfd::Block& b, int xi, bool load_tb) {
  const rfd::Xform& X = b.xf[xi];
  o << "apply_fixed<" << X.precalc << "u, " << ((X.flags & rfd::XF_HAS_POST) ? "true" : "false") << ", "
    << (load_tb ? std::string("TERM_BEGIN_LOAD") : std::to_string(X.term_begin) + "u") << ">(TermList<";
  for (uint32_t k = 0; k < X.n_pre; ++k) {
    const rfd::Term& T = b.terms[X.term_begin + k];
    o << (k ? ", " : "") << (T.var | (T.flags << 16)) << "u";
  }
  o << ">(), TermList<";
  for (uint32_t k = X.n_pre; k < X.n_terms; ++k) {
    const rfd::Term& T = b.terms[X.term_begin + k];
    o << (k > X.n_pre ? ", " : "") << (T.var | (T.flags << 16)) << "u";
  }
  o << ">(), blk, " << (xi == rfd::FINAL ? "FINAL" : std::to_string(xi));
}

}  // namespace

std::string structure_key(const rfd::Block& b) {
  std::ostringstream o;
  o << b.n_xforms << (b.has_final ? "F" : "") << ":";
  for (uint32_t i = 0; i < b.n_xforms; ++i) o << xform_key(b, (int)i) << ";";
  if (b.has_final) o << xform_key(b, rfd::FINAL) << ";";
  return o.str();
}

bool structure_covers(const rfd::Block& big, const rfd::Block& b) {
  if (b.n_xforms > big.n_xforms || (b.has_final != 0) != (big.has_final != 0)) return false;
  for (uint32_t i = 0; i < b.n_xforms; ++i)
    if (xform_key(b, (int)i) != xform_key(big, (int)i)) return false;
  return !b.has_final || xform_key(b, rfd::FINAL) == xform_key(big, rfd::FINAL);
}

std::string entry_name(int waves, int mode) {
  if (mode >= kBatchKernel) return "rf_batch_w" + std::to_string(waves) + "_h" + std::to_string(mode - kBatchKernel);
  return "rf_prog_w" + std::to_string(waves) + "_h" + std::to_string(mode);
}

std::string program_source(const rfd::Block& b, const std::vector<std::pair<int, int>>& kernels) {
  std::ostringstream o;
  o << "// flame program: " << structure_key(b) << "\n"
    << "typedef unsigned char uint8_t;\ntypedef unsigned short uint16_t;\ntypedef unsigned int uint32_t;\n"
    << "typedef unsigned long uint64_t;\ntypedef int int32_t;\ntypedef long int64_t;\n"
    << "#include \"rf_k_fast_body.hpp\"\n#include \"rf_k_batch_body.hpp\"\n"
    << "namespace rfprog {\nusing namespace rfd;\n"
    << "struct Prog {\n";
  const int n = (int)b.n_xforms;
  // flames run by this program have at most n xforms (see structure_covers)
  o << "  template <class BP>\n"
    << "  static __device__ __forceinline__ int select(BP blk, uint32_t raw) { return select_xform_upto<" << n
    << ">(blk, raw); }\n"
    << "  template <class Rng, class BP>\n"
    << "  static __device__ __forceinline__ XOut apply(BP blk, int xi, float x, float y, float c, bool first,\n"
    << "                                              Rng& rng) {\n"
    << "    switch (xi) {\n";
  for (int i = 0; i < n; ++i) {
    o << "      " << (i + 1 < n ? "case " + std::to_string(i) + ":" : std::string("default:")) << " return ";
    xform_args(o, b, i, false);
    o << ", x, y, c, first, rng);\n";
  }
  o << "    }\n  }\n"
    << "  template <class BP>\n"
    << "  static __device__ __forceinline__ bool has_final(BP) { return " << (b.has_final ? "true" : "false")
    << "; }\n";
  auto final_fn = [&](bool load_tb) {
    o << "  template <class Rng, class BP>\n"
      << "  static __device__ __forceinline__ XOut apply_final(BP blk, float x, float y, float c, Rng& rng) {\n";
    if (b.has_final) {
      o << "    return ";
      xform_args(o, b, rfd::FINAL, load_tb);
      o << ", x, y, c, false, rng);\n";
    } else {
      o << "    XOut r; r.x = x; r.y = y; r.c = c; r.w = 1.0f; return r;\n";
    }
    o << "  }\n";
  };
  final_fn(false);
  o << "};\n"
    << "// the batched renderer's view: the final xform's terms found through the Block\n"
    << "struct BProg : Prog {\n";
  final_fn(true);
  o << "};\n}  // namespace rfprog\n";
  for (const auto& k : kernels) {
    const int W = k.first, mode = k.second;
    o << "extern \"C\" __global__ __launch_bounds__(" << 64 * W << ") ";
    if (mode < kBatchKernel)
      o << "__attribute__((amdgpu_waves_per_eu(rfk::detail::fast_waves_per_eu<" << mode << ">()))) ";
    else if (mode == kBatchKernel + 5)
      o << "__attribute__((amdgpu_waves_per_eu(6))) ";
    o << "void " << entry_name(W, mode);
    if (mode >= kBatchKernel) {
      o << "(rfk::BatchArgs g, const rfd::Block* __restrict__ blocks,\n"
        << "    const rfk::BatchFrame* __restrict__ frames) {\n"
        << "  rfk::detail::k_batch_body<" << W << ", " << mode - kBatchKernel << ", " << lds_logk(W, mode)
        << ", rfprog::BProg>(g, blocks, frames);\n}\n";
    } else {
      o << "(rfk::FastArgs g, const rfd::Block* __restrict__ blocks) {\n"
        << "  rfk::detail::k_fast_body<" << W << ", " << mode << ", " << lds_logk(W, mode)
        << ", rfprog::Prog>(g, blocks);\n}\n";
    }
  }
  return o.str();
}

bool compile(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& log,
             double* seconds) {
  const auto t0 = std::chrono::steady_clock::now();
  constexpr int NH = (int)(sizeof(k_jit_headers) / sizeof(k_jit_headers[0]));
  const char* names[NH];
  const char* texts[NH];
  for (int i = 0; i < NH; ++i) { names[i] = k_jit_headers[i][0]; texts[i] = k_jit_headers[i][1]; }
  hiprtcProgram prog = nullptr;
  if (hiprtcCreateProgram(&prog, src.c_str(), "rf_flame_program.hip", NH, texts, names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  // The flags of the precompiled kernels (refrakt_amd/build.py): no implicit
  // contraction, GLSL-accuracy division and sqrt, no packed FP32 ops.
  const std::string arch_opt = "--offload-arch=" + arch;
  const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  while (!log.empty() && log.back() == '\0') log.pop_back();
  bool ok = r == HIPRTC_SUCCESS;
  if (ok) {
    size_t cs = 0;
    ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
    if (ok) {
      code.resize(cs);
      ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    }
  } else if (log.empty()) {
    log = hiprtcGetErrorString(r);
  }
  hiprtcDestroyProgram(&prog);
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return ok;
}

struct Program {
  std::string key;
  int device = -1;
  std::vector<hipModule_t> modules;
  std::map<std::pair<int, int>, hipFunction_t> fns;
  std::map<std::pair<int, int>, int> resident;  // persistent kernels: resident workgroups (grid size)
  double compile_s = 0.0;
};

namespace {
std::mutex g_mu;
std::map<std::string, std::unique_ptr<Program>> g_cache;  // key: device + structure

std::string device_arch(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "gfx950";
  std::string a = p.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}
}  // namespace

const Program* get_program(const rfd::Block& b, const std::vector<std::pair<int, int>>& kernels,
                           std::string& err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { err = "hipGetDevice failed"; return nullptr; }
  const std::string skey = structure_key(b);
  const std::string key = std::to_string(dev) + "/" + skey;
  std::lock_guard<std::mutex> lk(g_mu);
  auto& slot = g_cache[key];
  if (!slot) {
    slot.reset(new Program);
    slot->key = skey;
    slot->device = dev;
  }
  Program* p = slot.get();
  std::vector<std::pair<int, int>> missing;
  for (const auto& k : kernels)
    if (!p->fns.count(k) && std::find(missing.begin(), missing.end(), k) == missing.end()) missing.push_back(k);
  if (missing.empty()) return p;
  std::vector<char> code;
  std::string log;
  double sec = 0.0;
  if (!compile(program_source(b, missing), device_arch(dev), code, log, &sec)) {
    err = "flame program compilation failed: " + log;
    return nullptr;
  }
  p->compile_s += sec;
  hipModule_t mod = nullptr;
  hipError_t e = hipModuleLoadData(&mod, code.data());
  if (e != hipSuccess) { err = std::string("hipModuleLoadData: ") + hipGetErrorString(e); return nullptr; }
  p->modules.push_back(mod);
  for (const auto& k : missing) {
    hipFunction_t fn = nullptr;
    e = hipModuleGetFunction(&fn, mod, entry_name(k.first, k.second).c_str());
    if (e != hipSuccess) { err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e); return nullptr; }
    if (k.second == 4 || k.second >= kBatchKernel) {
      int per_cu = 0, cus = 0;
      e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * k.first, 0);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e != hipSuccess) { err = std::string("occupancy query: ") + hipGetErrorString(e); return nullptr; }
      p->resident[k] = cus * (per_cu > 0 ? per_cu : 1);
    }
    p->fns[k] = fn;
  }
  return p;
}

bool program_has(const Program* p, int device, const std::vector<std::pair<int, int>>& kernels) {
  if (!p || p->device != device) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  for (const auto& k : kernels)
    if (!p->fns.count(k)) return false;
  return true;
}

const Program* find_program(const rfd::Block& b, std::pair<int, int> kernel) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const std::string key = std::to_string(dev) + "/" + structure_key(b);
  std::lock_guard<std::mutex> lk(g_mu);
  const auto it = g_cache.find(key);
  if (it == g_cache.end() || !it->second || !it->second->fns.count(kernel)) return nullptr;
  return it->second.get();
}

double program_compile_seconds(const Program* p) { return p ? p->compile_s : 0.0; }

hipError_t launch_batch(const Program* p, const rfk::BatchArgs& a0, int waves, int hist_mode, uint32_t max_grid,
                        hipStream_t s) {
  const std::pair<int, int> k{waves, kBatchKernel + hist_mode};
  const auto it = p->fns.find(k);
  if (it == p->fns.end()) return hipErrorNotFound;
  uint32_t grid = hist_mode == 5 ? a0.n_items : std::min<uint32_t>(a0.n_items, (uint32_t)p->resident.at(k));
  if (max_grid) grid = std::min(grid, max_grid);
  if (grid == 0) return hipSuccess;
  rfk::BatchArgs a = a0;
  const rfd::Block* blocks = a.blocks;
  const rfk::BatchFrame* frames = a.frames;
  void* args[] = {&a, &blocks, &frames};
  return hipModuleLaunchKernel(it->second, grid, 1, 1, 64u * (uint32_t)waves, 1, 1, 0, s, args, nullptr);
}

hipError_t launch_fast(const Program* p, const rfk::FastArgs& a0, uint64_t n, int waves, int hist_mode,
                       hipStream_t s) {
  const auto it = p->fns.find({waves, hist_mode});
  if (it == p->fns.end()) return hipErrorNotFound;
  rfk::FastArgs a = a0;
  a.n_batches = (uint32_t)(n / (64u * (uint32_t)waves));
  uint32_t grid = a.n_batches;
  if (hist_mode == 4) grid = std::min<uint32_t>(grid, (uint32_t)p->resident.at({waves, hist_mode}));
  const rfd::Block* blocks = a.blocks;
  void* args[] = {&a, &blocks};
  return hipModuleLaunchKernel(it->second, grid, 1, 1, 64u * (uint32_t)waves, 1, 1, 0, s, args, nullptr);
}

}  // namespace rfj
