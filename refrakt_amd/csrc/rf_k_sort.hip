// Sort-based histogram accumulation: bucket -> tile split and LDS tile sums.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_kern_common.hpp"
#include "rf_sort.hpp"

namespace rfk {
namespace {
using namespace detail;

// Clears the cursors of a sort draw and, on a call's first sub-draw, the
// draw counters (one launch instead of a cursor reset plus a memset).
__global__ void k_sort_reset(SortArgs a, unsigned long long* __restrict__ counter, uint32_t counter_words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n_buckets) a.bcur[i] = 0u;
  if (i < a.n_tiles) a.tcur[i] = 0u;
  if (i < counter_words) counter[i] = 0ull;
}

// A draw's readback in one launch: the counter words, then the bucket and
// tile cursors, stored straight into pinned host memory (out is its device
// address), instead of one copy command per array.
__global__ void k_draw_readback(const unsigned long long* __restrict__ counter, uint32_t counter_words,
                                const uint32_t* __restrict__ bcur, uint32_t nb, const uint32_t* __restrict__ tcur,
                                uint32_t nt, uint32_t* __restrict__ out) {
  const uint32_t cw = 2u * counter_words, n = cw + nb + nt;
  const uint32_t* c32 = reinterpret_cast<const uint32_t*>(counter);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = i < cw ? c32[i] : (i < cw + nb ? bcur[i - cw] : tcur[i - cw - nb]);
}

// One workgroup per chunk of SORT_CHUNK records of a bucket (the host lists
// the chunks a bucket's capacity can hold): a counting sort of the chunk by
// tile in LDS, one reservation atomic per non-empty tile, then every record
// is written to its tile's run.  Records that do not fit in their tile's
// capacity are added to the histogram directly.  The tile of a sorted record
// is read back from the record itself, so the workgroup needs 69 KiB of LDS
// and two of them (32 waves) fit on a CU.  Larger chunks pay the per-chunk
// work (staging, reservations, scan, barriers) over more records: 8192 and
// 4096 records per workgroup took 12 % and 30 % longer on a 7680x4320 draw.
constexpr int RB_THREADS = 1024;
constexpr int RB_PER = SORT_CHUNK / RB_THREADS;  // records per thread

__global__ __launch_bounds__(RB_THREADS) void k_sort_rebucket(SortArgs a, float* __restrict__ bins,
                                                              const float4* __restrict__ palette) {
  static_assert(SORT_TILES_PER_BUCKET <= 4 * WAVE, "scan width");
  static_assert(SORT_TILES_PER_BUCKET <= RB_THREADS, "one thread per tile");
  static_assert(RB_PER % 4 == 0, "16-byte loads");
  __shared__ uint32_t s_out[SORT_CHUNK];  // tile-ordered records
  __shared__ uint32_t s_cnt[SORT_TILES_PER_BUCKET], s_off[SORT_TILES_PER_BUCKET], s_gb[SORT_TILES_PER_BUCKET];
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_toff[SORT_TILES_PER_BUCKET], s_tcap[SORT_TILES_PER_BUCKET];
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
  // every thread keeps its RB_PER records in registers between the passes;
  // a full chunk (all but a bucket's last) issues its loads without per-lane
  // guards, which would otherwise make each load wait for the one before
  uint32_t rec[RB_PER];
  if (n == (uint32_t)SORT_CHUNK) {  // workgroup-uniform
#pragma unroll
    for (int k = 0; k < RB_PER / 4; ++k) {
      const uint4 q = src4[(uint32_t)k * RB_THREADS + tid];
      rec[4 * k] = q.x; rec[4 * k + 1] = q.y; rec[4 * k + 2] = q.z; rec[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < RB_PER / 4; ++k) {
      const uint32_t i4 = (uint32_t)k * RB_THREADS + tid, i = 4u * i4;
      if (i + 3u < n) {
        const uint4 q = src4[i4];
        rec[4 * k] = q.x; rec[4 * k + 1] = q.y; rec[4 * k + 2] = q.z; rec[4 * k + 3] = q.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) rec[4 * k + e] = (i + (uint32_t)e) < n ? src[i + e] : SORT_NONE;
      }
    }
  }
  __syncthreads();
  constexpr uint32_t TSH = 8 + SORT_TILE_SHIFT;  // record >> TSH = tile within the bucket
  uint32_t rank[RB_PER];
#pragma unroll
  for (int m = 0; m < RB_PER; ++m) rank[m] = rec[m] != SORT_NONE ? atomicAdd(&s_cnt[rec[m] >> TSH], 1u) : 0u;
  __syncthreads();
  wave0_exclusive_scan<4>(s_cnt, s_off, SORT_TILES_PER_BUCKET, &s_total);
  if (tid < (uint32_t)SORT_TILES_PER_BUCKET) {
    const uint32_t c = s_cnt[tid];
    s_gb[tid] = c ? atomicAdd(a.tcur + b * SORT_TILES_PER_BUCKET + tid, c) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < RB_PER; ++m)
    if (rec[m] != SORT_NONE) s_out[s_off[rec[m] >> TSH] + rank[m]] = rec[m];
  __syncthreads();
  // Write-out, four rounds at a time: their LDS reads are issued before the
  // first store, and one ballot per four rounds skips the overflow path.
  const uint32_t total = s_total;
  constexpr int U = 4;
  for (uint32_t q0 = 0; q0 * RB_THREADS < total; q0 += U) {  // uniform
    uint32_t j[U], r[U], gb[U], off[U], cap[U], toff[U];
    bool in[U], ovf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      j[u] = (q0 + u) * RB_THREADS + tid;
      in[u] = j[u] < total;
      r[u] = s_out[in[u] ? j[u] : 0u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t f = r[u] >> TSH;
      gb[u] = s_gb[f];
      off[u] = s_off[f];
      cap[u] = s_tcap[f];
      toff[u] = s_toff[f];
    }
    bool any = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t dst = gb[u] + (j[u] - off[u]);
      ovf[u] = in[u] && dst >= cap[u];
      if (in[u] && !ovf[u])
        a.tbuf[toff[u] + dst] = (((r[u] >> 8) & ((1u << SORT_TILE_SHIFT) - 1u)) << 8) | (r[u] & 255u);
      any = any || ovf[u];
    }
    if (__builtin_amdgcn_ballot_w64(any)) {  // wave-uniform and rare
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t bin = ovf[u] ? (b << SORT_COARSE_SHIFT) | (r[u] >> 8) : 0u;
        // the tile's frame's palette (batched rendering), from global memory
        const float4* pal = a.tile_pal ? palette + a.tile_pal[bin >> SORT_TILE_SHIFT] : palette;
        bin_quad(bins, bin, ovf[u], r[u] & 255u, 1.0f, pal);
      }
    }
  }
}

// One workgroup per segment of a tile of 2^12 bins (the host lists the
// segments of every tile): the segment's records are summed in LDS with two
// 64-bit atomics per record, red | green << 32 and blue | count << 32, the
// colours in fixed point (see SORT_SEG16 and SORT_SEG15 for why no field
// carries into the next; a palette entry's three fields fit 16 bits each, so
// the table read per record is 8 bytes, which LDS serves in half the cycles
// of the 16-byte addends: 8 % less time per 7680x4320 draw), then the
// touched bins are added to the histogram:
// with a plain read-modify-write when the tile is one segment (tiles are
// disjoint, and every overflow was added by earlier kernels), with quad
// atomics when several segments share the tile.  68 KiB of LDS: two
// workgroups of 16 waves per CU (tiles of 2^11 bins, 36 KiB and four
// workgroups of 8 waves, took 30 % longer on a 1024x1024 draw: the per-tile
// clearing, palette and write-back are paid over half the records).  Each
// thread loads 16 records (four 16-byte loads in flight) before it issues
// their LDS atomics.
constexpr int TILE_THREADS = 1024;
constexpr int TILE_BINS = 1 << SORT_TILE_SHIFT;

struct alignas(8) PalQ {
  uint32_t rg;  // red | green << 16
  uint32_t b;   // blue
};
// the two 64-bit addends of a palette entry
__device__ __forceinline__ uint64_t pal_rg(const PalQ& q) {
  return (uint64_t)(q.rg & 0xffffu) | ((uint64_t)(q.rg >> 16) << 32);
}
__device__ __forceinline__ uint64_t pal_bc(const PalQ& q) { return (uint64_t)q.b | (1ull << 32); }
struct TileSums {
  uint64_t rg[TILE_BINS];  // red | green << 32
  uint64_t bc[TILE_BINS];  // blue | count << 32
  PalQ pal[256];               // per palette entry, one 16-byte LDS read
};

__device__ __forceinline__ void tile_add(TileSums& t, uint32_t r) {
  const uint32_t p = r & 255u, i = r >> 8;
  const PalQ q = t.pal[p];
  lds_add_u64(&t.rg[i], pal_rg(q));
  lds_add_u64(&t.bc[i], pal_bc(q));
}

// Adds K records: every palette read is issued before the first atomic, so a
// record's atomics do not each wait for their own read.
template <int K>
__device__ __forceinline__ void tile_add_batch(TileSums& t, const uint32_t* r) {
  PalQ v[K];
  uint32_t bi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    bi[k] = r[k] >> 8;
    v[k] = t.pal[r[k] & 255u];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    lds_add_u64(&t.rg[bi[k]], pal_rg(v[k]));
    lds_add_u64(&t.bc[bi[k]], pal_bc(v[k]));
  }
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
  // workgroup-uniform: a tile of at most SORT_SEG16 records is one segment
  // summed at 2^16 - 1 per unit, a larger one segments of SORT_SEG15 at
  // 2^15 - 1
  const bool split = cap > (uint32_t)SORT_SEG16;
  const uint32_t seg = split ? (uint32_t)SORT_SEG15 : (uint32_t)SORT_SEG16;
  const float one = split ? 32767.0f : 65535.0f;
  const uint32_t start = (job & 0xffffu) * seg;
  if (base >= n_bins || start >= count) return;  // workgroup-uniform
  const uint32_t n = (count - start) < seg ? (count - start) : seg;
  for (uint32_t i = tid; i < (uint32_t)TILE_BINS; i += TILE_THREADS) {
    s_t.rg[i] = 0ull;
    s_t.bc[i] = 0ull;
  }
  if (tid < 256u) {
    // palette values lie in [0, 1]: each field gets at most `one` (< 2^16)
    // per record
    const float4 c = (a.tile_pal ? palette + a.tile_pal[t] : palette)[tid];
    const uint32_t r = (uint32_t)(c.x * one + 0.5f), g = (uint32_t)(c.y * one + 0.5f),
                   bl = (uint32_t)(c.z * one + 0.5f);
    s_t.pal[tid].rg = r | (g << 16);
    s_t.pal[tid].b = bl;
  }
  __syncthreads();
  // Tile offsets and both segment sizes are multiples of 4 records, so the
  // segment starts 16-byte aligned.
  const uint32_t* src = a.tbuf + a.toff[t] + start;
  const uint4* src4 = reinterpret_cast<const uint4*>(src);
  const uint32_t n4 = n / 4;
  constexpr uint32_t G = 4;  // 16-byte loads per thread and round
  for (uint32_t i0 = 0; i0 < n4; i0 += G * TILE_THREADS) {
    if (i0 + G * TILE_THREADS <= n4) {  // workgroup-uniform: a full round
      uint32_t r[4 * G];
#pragma unroll
      for (uint32_t k = 0; k < G; ++k) {
        const uint4 q = src4[i0 + k * TILE_THREADS + tid];
        r[4 * k] = q.x; r[4 * k + 1] = q.y; r[4 * k + 2] = q.z; r[4 * k + 3] = q.w;
      }
      tile_add_batch<2 * G>(s_t, r);
      tile_add_batch<2 * G>(s_t, r + 2 * G);
    } else {
#pragma unroll
      for (uint32_t k = 0; k < G; ++k) {
        const uint32_t i = i0 + k * TILE_THREADS + tid;
        if (i < n4) {
          const uint4 q = src4[i];
          const uint32_t r[4] = {q.x, q.y, q.z, q.w};
          tile_add_batch<4>(s_t, r);
        }
      }
    }
  }
  for (uint32_t i = n4 * 4 + tid; i < n; i += TILE_THREADS) tile_add(s_t, src[i]);
  __syncthreads();
  const bool shared = cap > seg;  // several segments add to this tile
  const float inv = 1.0f / one;
  constexpr int BPT = TILE_BINS / TILE_THREADS;  // bins per thread
  static_assert(BPT * TILE_THREADS == TILE_BINS, "bins per thread");
  float4 v[BPT];
  bool add[BPT];
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    const uint32_t i = (uint32_t)k * TILE_THREADS + tid;
    const uint64_t bc = s_t.bc[i];
    const uint64_t rg = s_t.rg[i];
    const uint32_t c = (uint32_t)(bc >> 32);
    v[k] = make_float4((float)(uint32_t)rg * inv, (float)(uint32_t)(rg >> 32) * inv, (float)(uint32_t)bc * inv,
                       (float)c);
    add[k] = base + i < n_bins && c != 0u;
  }
  if (shared) {  // workgroup-uniform
#pragma unroll
    for (int k = 0; k < BPT; ++k)
      bin_quad_vals(reinterpret_cast<float*>(hist), base + (uint32_t)k * TILE_THREADS + tid, add[k], v[k]);
  } else {
    // the touched bins' read-modify-writes, all loads in flight before the
    // first store
    float4 h[BPT];
#pragma unroll
    for (int k = 0; k < BPT; ++k)
      h[k] = add[k] ? hist[base + (uint32_t)k * TILE_THREADS + tid] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
      if (!add[k]) continue;
      h[k].x += v[k].x; h[k].y += v[k].y; h[k].z += v[k].z; h[k].w += v[k].w;
      hist[base + (uint32_t)k * TILE_THREADS + tid] = h[k];
    }
  }
}

// One workgroup per tile: the sum of the count channel over its bins, the
// estimate of where the next draw's samples go when a flame has no sort
// draw to learn from yet.
__global__ __launch_bounds__(256) void k_sort_tile_mass(const float4* __restrict__ hist, uint32_t n_bins,
                                                       float* __restrict__ mass) {
  __shared__ float s_w[256 / WAVE];
  const uint32_t base = blockIdx.x * (uint32_t)TILE_BINS;
  float sum = 0.0f;
  for (uint32_t i = threadIdx.x; i < (uint32_t)TILE_BINS; i += 256u)
    if (base + i < n_bins) sum += hist[base + i].w;
  sum = wave_sum(sum);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_w[threadIdx.x / WAVE] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int k = 0; k < 256 / WAVE; ++k) t += s_w[k];
    mass[blockIdx.x] = t;
  }
}

}  // namespace

hipError_t launch_sort_tile_mass(const float4* bins, uint32_t n_bins, uint32_t n_tiles, float* mass, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_tile_mass, dim3(n_tiles), dim3(256), 0, s, bins, n_bins, mass);
  return hipGetLastError();
}

hipError_t launch_sort_reset(const SortArgs& a, hipStream_t s, unsigned long long* counter, uint32_t counter_words) {
  uint32_t n = a.n_tiles > a.n_buckets ? a.n_tiles : a.n_buckets;
  if (!counter) counter_words = 0;
  if (counter_words > n) n = counter_words;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_reset, dim3((n + 255) / 256), dim3(256), 0, s, a, counter, counter_words);
  return hipGetLastError();
}

hipError_t launch_draw_readback(const unsigned long long* counter, uint32_t counter_words, const uint32_t* bcur,
                                uint32_t nb, const uint32_t* tcur, uint32_t nt, uint32_t* out, hipStream_t s) {
  const uint32_t n = 2u * counter_words + (bcur ? nb : 0u) + (tcur ? nt : 0u);
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 255) / 256 < 64u ? (n + 255) / 256 : 64u;
  hipLaunchKernelGGL(k_draw_readback, dim3(blocks), dim3(256), 0, s, counter, counter_words, bcur, bcur ? nb : 0u,
                     tcur, tcur ? nt : 0u, out);
  return hipGetLastError();
}

hipError_t launch_sort_rebucket(const SortArgs& a, float* bins, const float4* palette, hipStream_t s) {
  if (a.n_chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_rebucket, dim3(a.n_chunks), dim3(RB_THREADS), 0, s, a, bins, palette);
  return hipGetLastError();
}

hipError_t launch_sort_tiles(const SortArgs& a, float* bins, uint32_t n_bins, const float4* palette, hipStream_t s) {
  if (a.n_segs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_tiles, dim3(a.n_segs), dim3(TILE_THREADS), 0, s, a, reinterpret_cast<float4*>(bins),
                     n_bins, palette);
  return hipGetLastError();
}

}  // namespace rfk
