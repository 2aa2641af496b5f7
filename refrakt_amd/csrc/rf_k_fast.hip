// Fast engine: persistent chaos game with wave-uniform xforms and an LDS particle exchange.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rf_k_fast_body.hpp"

using namespace rfd;

namespace rfk {
namespace {
using namespace detail;

template <int WAVES, int HM, int LOGK>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(fast_waves_per_eu<HM>())))
void k_fast(FastArgs g, const Block* __restrict__ blocks) {
  k_fast_body<WAVES, HM, LOGK, InterpProg>(g, blocks);
}

}  // namespace

template <int W, int HM, int LOGK>
static hipError_t launch_fast_k(const FastArgs& a0, uint64_t n, hipStream_t s) {
  FastArgs a = a0;
  a.n_batches = (uint32_t)(n / (64u * W));
  uint32_t grid = a.n_batches;
  if (HM == 4) {
    // one workgroup per resident slot (cached per instantiation and device),
    // so that each LDS table spans many batches before it is flushed
    constexpr int kMaxDev = 64;
    static int resident[kMaxDev] = {0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int res = dev < kMaxDev ? resident[dev] : 0;
    if (!res) {
      int cus = 0, per_cu = 0;
      e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_fast<W, HM, LOGK>),
                                                         64 * W, 0);
      if (e != hipSuccess) return e;
      res = cus * (per_cu > 0 ? per_cu : 1);
      if (dev < kMaxDev) resident[dev] = res;
    }
    grid = std::min<uint32_t>(grid, (uint32_t)res);
  }
  hipLaunchKernelGGL((k_fast<W, HM, LOGK>), dim3(grid), dim3(64 * W), 0, s, a, a.blocks);
  return hipGetLastError();
}

template <int W>
static hipError_t launch_fast_w(const FastArgs& a, uint64_t n, int hm, hipStream_t s) {
  switch (hm) {
    case 2: return launch_fast_k<W, 2, 0>(a, n, s);
    case 3: return launch_fast_k<W, 3, 0>(a, n, s);
    case 4: return launch_fast_k<W, 4, 10>(a, n, s);
    case 5: return launch_fast_k<W, 5, 0>(a, n, s);
    default: return launch_fast_k<W, 1, 0>(a, n, s);
  }
}

hipError_t launch_fast(const FastArgs& a, uint64_t n_particles, int waves, int hist_mode, hipStream_t s) {
  switch (waves) {
    case 8: return launch_fast_w<8>(a, n_particles, hist_mode, s);
    default: return launch_fast_w<4>(a, n_particles, hist_mode, s);
  }
}

}  // namespace rfk
