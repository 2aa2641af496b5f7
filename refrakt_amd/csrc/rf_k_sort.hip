// Sort-based histogram accumulation: bucket -> tile split and LDS tile sums.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_kern_common.hpp"
#include "rf_sort.hpp"

namespace rfk {
namespace {
using namespace detail;

__global__ void k_sort_reset(SortArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n_buckets) a.bcur[i] = 0u;
  if (i < a.n_tiles) a.tcur[i] = 0u;
}

// One workgroup per chunk of SORT_CHUNK records of a bucket (the host lists
// the chunks a bucket's capacity can hold): a counting sort of the chunk by
// tile in LDS, one reservation atomic per non-empty tile, then every record
// is written to its tile's run.  Records that do not fit in their tile's
// capacity are added to the histogram directly.
constexpr int RB_PER = SORT_CHUNK / 256;  // records per thread

__global__ __launch_bounds__(256) void k_sort_rebucket(SortArgs a, float* __restrict__ bins,
                                                       const float4* __restrict__ palette) {
  static_assert(SORT_TILES_PER_BUCKET <= 4 * WAVE, "scan width");
  static_assert(SORT_TILES_PER_BUCKET <= 256, "tile index in a byte");
  static_assert(RB_PER % 4 == 0, "16-byte loads");
  __shared__ uint32_t s_out[SORT_CHUNK];   // tile-ordered records
  __shared__ uint8_t s_tile[SORT_CHUNK];   // tile (within the bucket) of s_out[j]
  __shared__ uint32_t s_cnt[SORT_TILES_PER_BUCKET], s_off[SORT_TILES_PER_BUCKET], s_gb[SORT_TILES_PER_BUCKET];
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_toff[SORT_TILES_PER_BUCKET], s_tcap[SORT_TILES_PER_BUCKET];
  __shared__ float4 s_pal[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t job = a.chunks[blockIdx.x];
  const uint32_t b = job >> 16;
  const uint32_t cap = a.bcap[b];
  const uint32_t cur = a.bcur[b];
  const uint32_t count = cur < cap ? cur : cap;
  const uint32_t start = (job & 0xffffu) * (uint32_t)SORT_CHUNK;
  if (start >= count) return;  // workgroup-uniform
  const uint32_t n = (count - start) < (uint32_t)SORT_CHUNK ? (count - start) : (uint32_t)SORT_CHUNK;
  // bucket offsets and SORT_CHUNK are multiples of 4 records: 16-byte aligned
  const uint32_t* src = a.cbuf + a.boff[b] + start;
  const uint4* src4 = reinterpret_cast<const uint4*>(src);
  if (tid < (uint32_t)SORT_TILES_PER_BUCKET) {
    s_cnt[tid] = 0u;
    s_toff[tid] = a.toff[b * SORT_TILES_PER_BUCKET + tid];
    s_tcap[tid] = a.tcap[b * SORT_TILES_PER_BUCKET + tid];
  }
  s_pal[tid] = palette[tid];
  // every thread keeps its RB_PER records in registers between the passes
  uint32_t rec[RB_PER];
#pragma unroll
  for (int k = 0; k < RB_PER / 4; ++k) {
    const uint32_t i4 = (uint32_t)k * 256u + tid, i = 4u * i4;
    if (i + 3u < n) {
      const uint4 q = src4[i4];
      rec[4 * k] = q.x; rec[4 * k + 1] = q.y; rec[4 * k + 2] = q.z; rec[4 * k + 3] = q.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) rec[4 * k + e] = (i + (uint32_t)e) < n ? src[i + e] : SORT_NONE;
    }
  }
  __syncthreads();
  uint32_t rank[RB_PER];
#pragma unroll
  for (int m = 0; m < RB_PER; ++m)
    rank[m] = rec[m] != SORT_NONE ? atomicAdd(&s_cnt[rec[m] >> (8 + SORT_TILE_SHIFT)], 1u) : 0u;
  __syncthreads();
  wave0_exclusive_scan<4>(s_cnt, s_off, SORT_TILES_PER_BUCKET, &s_total);
  if (tid < (uint32_t)SORT_TILES_PER_BUCKET) {
    const uint32_t c = s_cnt[tid];
    s_gb[tid] = c ? atomicAdd(a.tcur + b * SORT_TILES_PER_BUCKET + tid, c) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < RB_PER; ++m) {
    if (rec[m] != SORT_NONE) {
      const uint32_t f = rec[m] >> (8 + SORT_TILE_SHIFT);
      const uint32_t j = s_off[f] + rank[m];
      s_out[j] = rec[m];
      s_tile[j] = (uint8_t)f;
    }
  }
  __syncthreads();
  const uint32_t total = s_total;
  const uint32_t rounds = (total + 255u) / 256u;  // uniform
  for (uint32_t q = 0; q < rounds; ++q) {
    const uint32_t j = q * 256u + tid;
    bool ovf = false;
    uint32_t bin = 0, pidx = 0;
    if (j < total) {
      const uint32_t f = s_tile[j];
      const uint32_t r = s_out[j];
      const uint32_t local = r >> 8;
      const uint32_t dst = s_gb[f] + (j - s_off[f]);
      if (dst < s_tcap[f]) {
        a.tbuf[s_toff[f] + dst] = ((local & ((1u << SORT_TILE_SHIFT) - 1u)) << 8) | (r & 255u);
      } else {
        ovf = true;
        bin = (b << SORT_COARSE_SHIFT) | local;
        pidx = r & 255u;
      }
    }
    bin_quad(bins, bin, ovf, pidx, 1.0f, s_pal);
  }
}

// One workgroup per segment of up to SORT_SEG records of a tile of 2^11 bins
// (the host splits each tile's capacity into segments, so a hot tile is
// summed by many CUs): the segment's records are summed in LDS in fixed
// point (u64 colour planes, u32 counts: integer LDS atomics run at full rate),
// then the non-empty bins are added to the histogram: with a plain
// read-modify-write when the tile has a single segment (tiles are disjoint,
// and every overflow was added by earlier kernels), with quad atomics when
// several segments share the tile.  Each thread loads 16 records (four
// 16-byte loads in flight) before it issues their LDS atomics.
constexpr int TILE_THREADS = 512;
constexpr int TILE_BINS = 1 << SORT_TILE_SHIFT;

struct TileSums {
  uint64_t rgb[3][TILE_BINS];
  uint32_t count[TILE_BINS];
  uint32_t palq[3][256];
};

__device__ __forceinline__ void tile_add(TileSums& t, uint32_t r) {
  const uint32_t p = r & 255u, i = r >> 8;
  lds_add_u64(&t.rgb[0][i], (uint64_t)t.palq[0][p]);
  lds_add_u64(&t.rgb[1][i], (uint64_t)t.palq[1][p]);
  lds_add_u64(&t.rgb[2][i], (uint64_t)t.palq[2][p]);
  __hip_atomic_fetch_add(&t.count[i], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ __launch_bounds__(TILE_THREADS) void k_sort_tiles(SortArgs a, float4* __restrict__ hist, uint32_t n_bins,
                                                            const float4* __restrict__ palette) {
  __shared__ TileSums s_t;
  const uint32_t tid = threadIdx.x;
  const uint32_t job = a.segs[blockIdx.x];
  const uint32_t t = job >> 16;
  const uint32_t base = t << SORT_TILE_SHIFT;
  const uint32_t cap = a.tcap[t];
  const uint32_t cur = a.tcur[t];
  const uint32_t count = cur < cap ? cur : cap;
  const uint32_t start = (job & 0xffffu) * (uint32_t)SORT_SEG;
  if (base >= n_bins || start >= count) return;  // workgroup-uniform
  const uint32_t end = (count - start) < (uint32_t)SORT_SEG ? count : start + (uint32_t)SORT_SEG;
  for (uint32_t i = tid; i < (uint32_t)TILE_BINS; i += TILE_THREADS) {
    s_t.rgb[0][i] = 0ull;
    s_t.rgb[1][i] = 0ull;
    s_t.rgb[2][i] = 0ull;
    s_t.count[i] = 0u;
  }
  if (tid < 256u) {
    const float4 c = palette[tid];
    s_t.palq[0][tid] = (uint32_t)to_fix(c.x);
    s_t.palq[1][tid] = (uint32_t)to_fix(c.y);
    s_t.palq[2][tid] = (uint32_t)to_fix(c.z);
  }
  __syncthreads();
  // Tile offsets and SORT_SEG are multiples of 4 records, so the segment
  // starts 16-byte aligned.
  const uint32_t* src = a.tbuf + a.toff[t] + start;
  const uint4* src4 = reinterpret_cast<const uint4*>(src);
  const uint32_t n = end - start;
  const uint32_t n4 = n / 4;
  constexpr uint32_t G = 4;  // 16-byte loads per thread and round
  for (uint32_t i0 = 0; i0 < n4; i0 += G * TILE_THREADS) {
    uint4 q[G];
#pragma unroll
    for (uint32_t k = 0; k < G; ++k) {
      const uint32_t i = i0 + k * TILE_THREADS + tid;
      q[k] = i < n4 ? src4[i] : make_uint4(SORT_NONE, SORT_NONE, SORT_NONE, SORT_NONE);
    }
#pragma unroll
    for (uint32_t k = 0; k < G; ++k) {
      if (q[k].x != SORT_NONE) tile_add(s_t, q[k].x);
      if (q[k].y != SORT_NONE) tile_add(s_t, q[k].y);
      if (q[k].z != SORT_NONE) tile_add(s_t, q[k].z);
      if (q[k].w != SORT_NONE) tile_add(s_t, q[k].w);
    }
  }
  for (uint32_t i = n4 * 4 + tid; i < n; i += TILE_THREADS) tile_add(s_t, src[i]);
  __syncthreads();
  const bool shared = cap > (uint32_t)SORT_SEG;  // several segments may add to this tile
  for (uint32_t i = tid; i < (uint32_t)TILE_BINS; i += TILE_THREADS) {  // uniform trip count
    const uint32_t bin = base + i;
    const uint32_t c = s_t.count[i];
    const float4 s = make_float4(from_fix(s_t.rgb[0][i]), from_fix(s_t.rgb[1][i]), from_fix(s_t.rgb[2][i]),
                                 (float)c);
    const bool add = bin < n_bins && c != 0u;
    if (shared) {
      bin_quad_vals(reinterpret_cast<float*>(hist), bin, add, s);
    } else if (add) {
      float4 h = hist[bin];
      h.x += s.x; h.y += s.y; h.z += s.z; h.w += s.w;
      hist[bin] = h;
    }
  }
}

// Tiles of at most SORT_SEG_PACKED records: one workgroup per tile, two
// 64-bit LDS atomics per record (see SORT_SEG_PACKED), then one plain
// read-modify-write per touched bin.  36 KiB of LDS: four workgroups per CU.
struct TileSumsPacked {
  uint64_t rg[TILE_BINS];   // red | green << 32, 2^16 per unit
  uint64_t bc[TILE_BINS];   // blue | count << 32
  uint64_t prg[256];        // per palette entry: red | green << 32
  uint64_t pb[256];         // blue | 1 << 32
};

__device__ __forceinline__ uint32_t fix16(float v) { return (uint32_t)(v * 65536.0f + 0.5f); }

__device__ __forceinline__ void tile_add_packed(TileSumsPacked& t, uint32_t r) {
  const uint32_t p = r & 255u, i = r >> 8;
  lds_add_u64(&t.rg[i], t.prg[p]);
  lds_add_u64(&t.bc[i], t.pb[p]);
}

__global__ __launch_bounds__(TILE_THREADS) void k_sort_tiles_packed(SortArgs a, float4* __restrict__ hist,
                                                                   uint32_t n_bins,
                                                                   const float4* __restrict__ palette) {
  __shared__ TileSumsPacked s_t;
  const uint32_t tid = threadIdx.x;
  const uint32_t t = a.psegs[blockIdx.x];
  const uint32_t base = t << SORT_TILE_SHIFT;
  const uint32_t cap = a.tcap[t];
  const uint32_t cur = a.tcur[t];
  const uint32_t n = cur < cap ? cur : cap;  // <= SORT_SEG_PACKED: every field stays below 2^32
  if (base >= n_bins || n == 0) return;     // workgroup-uniform
  for (uint32_t i = tid; i < (uint32_t)TILE_BINS; i += TILE_THREADS) {
    s_t.rg[i] = 0ull;
    s_t.bc[i] = 0ull;
  }
  if (tid < 256u) {
    const float4 c = palette[tid];
    s_t.prg[tid] = (uint64_t)fix16(c.x) | ((uint64_t)fix16(c.y) << 32);
    s_t.pb[tid] = (uint64_t)fix16(c.z) | (1ull << 32);
  }
  __syncthreads();
  const uint32_t* src = a.tbuf + a.toff[t];  // 16-byte aligned (tile offsets are multiples of 4)
  const uint4* src4 = reinterpret_cast<const uint4*>(src);
  const uint32_t n4 = n / 4;
  constexpr uint32_t G = 4;
  for (uint32_t i0 = 0; i0 < n4; i0 += G * TILE_THREADS) {
    uint4 q[G];
#pragma unroll
    for (uint32_t k = 0; k < G; ++k) {
      const uint32_t i = i0 + k * TILE_THREADS + tid;
      q[k] = i < n4 ? src4[i] : make_uint4(SORT_NONE, SORT_NONE, SORT_NONE, SORT_NONE);
    }
#pragma unroll
    for (uint32_t k = 0; k < G; ++k) {
      if (q[k].x != SORT_NONE) tile_add_packed(s_t, q[k].x);
      if (q[k].y != SORT_NONE) tile_add_packed(s_t, q[k].y);
      if (q[k].z != SORT_NONE) tile_add_packed(s_t, q[k].z);
      if (q[k].w != SORT_NONE) tile_add_packed(s_t, q[k].w);
    }
  }
  for (uint32_t i = n4 * 4 + tid; i < n; i += TILE_THREADS) tile_add_packed(s_t, src[i]);
  __syncthreads();
  constexpr float INV = 1.0f / 65536.0f;
  for (uint32_t i = tid; i < (uint32_t)TILE_BINS; i += TILE_THREADS) {
    const uint32_t bin = base + i;
    const uint64_t bc = s_t.bc[i];
    const uint32_t c = (uint32_t)(bc >> 32);
    if (bin < n_bins && c != 0u) {
      const uint64_t rg = s_t.rg[i];
      float4 h = hist[bin];
      h.x += (float)(uint32_t)rg * INV;
      h.y += (float)(uint32_t)(rg >> 32) * INV;
      h.z += (float)(uint32_t)bc * INV;
      h.w += (float)c;
      hist[bin] = h;
    }
  }
}

}  // namespace

hipError_t launch_sort_reset(const SortArgs& a, hipStream_t s) {
  const uint32_t n = a.n_tiles > a.n_buckets ? a.n_tiles : a.n_buckets;
  hipLaunchKernelGGL(k_sort_reset, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sort_rebucket(const SortArgs& a, float* bins, const float4* palette, hipStream_t s) {
  if (a.n_chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_rebucket, dim3(a.n_chunks), dim3(256), 0, s, a, bins, palette);
  return hipGetLastError();
}

hipError_t launch_sort_tiles(const SortArgs& a, float* bins, uint32_t n_bins, const float4* palette, hipStream_t s) {
  if (a.n_psegs) {
    hipLaunchKernelGGL(k_sort_tiles_packed, dim3(a.n_psegs), dim3(TILE_THREADS), 0, s, a,
                       reinterpret_cast<float4*>(bins), n_bins, palette);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (a.n_segs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_tiles, dim3(a.n_segs), dim3(TILE_THREADS), 0, s, a, reinterpret_cast<float4*>(bins),
                     n_bins, palette);
  return hipGetLastError();
}

}  // namespace rfk
