// Stable LSD radix sort of (key, uint32 value) pairs for gfx950, 8 bits per pass.
//
// Used for the Morton orders of points and queries and for the CSR transpose of the
// neighbour table.  Everything is stream-ordered kernels over caller-provided
// workspace: no allocation, no memset, no host synchronisation, so the whole sort can
// be captured in a graph.  Stable and deterministic: equal keys keep input order.
//
// Per pass over digit (key >> shift) & 255, with tiles of kTileKeys keys per block:
//   1. digit_hist: per-tile digit counts (LDS), stored digit-major: counts[d * tiles + t]
//   2. exclusive scan of counts (3 kernels): start of each (digit, tile) in the output
//   3. digit_scatter: local digit starts from the tile's own counts (kept by step 1);
//      each tile walks its keys in order, 256 at a time; a key's rank
//      among equal digits is (earlier chunks) + (earlier waves of this chunk) +
//      (earlier lanes of its wave, from 8 ballots over the digit bits).  Keys are
//      placed in LDS at their tile-local rank, then written out run by run.
#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kTileKeys = kSortThreads * kSortItems;  // 4096
constexpr int kScanChunk = 4096;                      // elements per scan block

template <typename K>
__global__ __launch_bounds__(kSortThreads) void digit_hist_kernel(const K* __restrict__ keys, int64_t n,
                                                                  int shift, int64_t tiles,
                                                                  uint32_t* __restrict__ counts,
                                                                  uint32_t* __restrict__ tile_counts) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTileKeys;
#pragma unroll 4
  for (int i = 0; i < kSortItems; ++i) {
    const int64_t j = base + i * kSortThreads + threadIdx.x;
    if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);  // integer: order-independent
  }
  __syncthreads();
  counts[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];       // scanned next
  tile_counts[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x] = h[threadIdx.x];  // kept per tile
}

// ---- exclusive scan of uint32 (in place), three kernels
__global__ __launch_bounds__(256) void scan_reduce_kernel(const uint32_t* __restrict__ x, int64_t n,
                                                          uint32_t* __restrict__ sums) {
  __shared__ uint32_t s[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  uint32_t acc = 0;
  for (int i = threadIdx.x; i < kScanChunk; i += 256) {
    const int64_t j = base + i;
    if (j < n) acc += x[j];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// Block-wide exclusive scan of 256 values (Hillis-Steele in LDS); returns the total.
__device__ __forceinline__ uint32_t block_exscan256(uint32_t* s, uint32_t v, uint32_t* out) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t t = static_cast<int>(threadIdx.x) >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  const uint32_t incl = s[threadIdx.x];
  const uint32_t total = s[255];
  __syncthreads();
  *out = incl - v;
  return total;
}

__global__ __launch_bounds__(256) void scan_top_kernel(uint32_t* __restrict__ sums, int64_t m) {
  __shared__ uint32_t s[256];
  uint32_t carry = 0;
  for (int64_t base = 0; base < m; base += 256) {
    const int64_t j = base + threadIdx.x;
    const uint32_t v = j < m ? sums[j] : 0u;
    uint32_t ex;
    const uint32_t total = block_exscan256(s, v, &ex);
    if (j < m) sums[j] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void scan_down_kernel(uint32_t* __restrict__ x, int64_t n,
                                                        const uint32_t* __restrict__ sums) {
  __shared__ uint32_t s[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  constexpr int per = kScanChunk / 256;  // consecutive elements per thread
  uint32_t v[per];
  uint32_t local = 0;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    const int64_t j = base + threadIdx.x * per + i;
    v[i] = j < n ? x[j] : 0u;
    local += v[i];
  }
  uint32_t ex;
  block_exscan256(s, local, &ex);
  uint32_t run = sums[blockIdx.x] + ex;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    const int64_t j = base + threadIdx.x * per + i;
    if (j < n) x[j] = run;
    run += v[i];
  }
}

// Small scans (the sorts of up to 128K keys): one block of 1024 threads, each thread
// owning a run of consecutive elements -- one launch instead of three.
constexpr int kScanOneThreads = 1024;
constexpr int64_t kScanOneMax = 8192;

__global__ __launch_bounds__(kScanOneThreads) void scan_one_block_kernel(uint32_t* __restrict__ x, int64_t n) {
  __shared__ uint32_t wsum[kScanOneThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (n + kScanOneThreads - 1) / kScanOneThreads;
  const int64_t b = threadIdx.x * per;
  const int64_t e = b + per < n ? b + per : n;
  uint32_t local = 0;
  for (int64_t j = b; j < e; ++j) local += x[j];
  // inclusive scan within the wave, then over the wave totals
  uint32_t incl = local;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  if (wv == 0) {
    uint32_t w = lane < kScanOneThreads / 64 ? wsum[lane] : 0u;
#pragma unroll
    for (int off = 1; off < kScanOneThreads / 64; off <<= 1) {
      const uint32_t t = __shfl_up(w, off, 64);
      if (lane >= off) w += t;
    }
    if (lane < kScanOneThreads / 64) wsum[lane] = w;  // inclusive over waves
  }
  __syncthreads();
  uint32_t run = (wv > 0 ? wsum[wv - 1] : 0u) + incl - local;
  for (int64_t j = b; j < e; ++j) {
    const uint32_t v = x[j];
    x[j] = run;
    run += v;
  }
}

static hipError_t exclusive_scan(uint32_t* x, int64_t n, uint32_t* sums, hipStream_t st) {
  if (n <= kScanOneMax) {
    hipLaunchKernelGGL(scan_one_block_kernel, dim3(1), dim3(kScanOneThreads), 0, st, x, n);
    return hipGetLastError();
  }
  const int64_t chunks = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(chunks)), dim3(256), 0, st, x, n, sums);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, st, sums, chunks);
  hipLaunchKernelGGL(scan_down_kernel, dim3(static_cast<unsigned>(chunks)), dim3(256), 0, st, x, n, sums);
  return hipGetLastError();
}

// ---- stable scatter through LDS
// The tile is ranked stably into LDS (digit-major, input order within a digit), then
// written out in LDS order: consecutive threads write consecutive addresses of one
// digit's run, so the global stores coalesce.
template <typename K>
__global__ __launch_bounds__(kSortThreads) void digit_scatter_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift, int64_t tiles,
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ tile_counts, K* __restrict__ kout,
    uint32_t* __restrict__ vout) {
  __shared__ K skey[kTileKeys];
  __shared__ uint32_t sval[kTileKeys];
  __shared__ uint32_t loc[256];        // local start of each digit in the tile
  __shared__ uint32_t gpos[256];       // global start of each digit's run of this tile
  __shared__ uint32_t run[256];        // next local slot of each digit
  __shared__ uint32_t wcnt[4][256];    // per-wave digit counts of the current chunk
  __shared__ uint32_t scan_tmp[256];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTileKeys;
  const int tile_n = static_cast<int>(min<int64_t>(kTileKeys, n - base));

  // local digit starts from the tile's histogram (digit_hist), global starts from the scan
  uint32_t ex;
  block_exscan256(scan_tmp, tile_counts[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x], &ex);
  loc[threadIdx.x] = ex;
  run[threadIdx.x] = ex;
  gpos[threadIdx.x] = start[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x];
  __syncthreads();

  // all of this thread's keys and values are loaded up front, so their loads are in
  // flight together rather than one chunk at a time behind the barriers
  K kreg[kSortItems];
  uint32_t vreg[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int jj = i * kSortThreads + threadIdx.x;
    const bool ok = jj < tile_n;
    kreg[i] = ok ? kin[base + jj] : K(0);
    vreg[i] = ok ? (vin ? vin[base + jj] : static_cast<uint32_t>(base + jj)) : 0u;
  }
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
#pragma unroll
    for (int w = 0; w < 4; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const int jj = i * kSortThreads + threadIdx.x;
    const bool ok = jj < tile_n;
    const K key = kreg[i];
    const uint32_t val = vreg[i];
    const uint32_t d = ok ? static_cast<uint32_t>((key >> shift) & 255u) : 256u;
    unsigned long long same = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long m = __ballot(ok && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank_w = static_cast<uint32_t>(__popcll(same & lt));
    if (ok && rank_w == 0) wcnt[wave][d] = static_cast<uint32_t>(__popcll(same));
    __syncthreads();
    if (ok) {
      uint32_t off = run[d] + rank_w;
      for (int w = 0; w < wave; ++w) off += wcnt[w][d];
      skey[off] = key;
      sval[off] = val;
    }
    __syncthreads();
    run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
    __syncthreads();
  }
  // coalesced write-out of the locally sorted tile
  for (int jj = threadIdx.x; jj < tile_n; jj += kSortThreads) {
    const K key = skey[jj];
    const uint32_t d = static_cast<uint32_t>((key >> shift) & 255u);
    const uint32_t pos = gpos[d] + (jj - loc[d]);
    kout[pos] = key;
    vout[pos] = sval[jj];
  }
}

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

size_t radix_sort_workspace(int64_t n, size_t key_bytes) {
  const int64_t tiles = (n + kTileKeys - 1) / kTileKeys;
  const int64_t m = tiles * 256;
  const int64_t chunks = (m + kScanChunk - 1) / kScanChunk;
  return al(n * key_bytes) + al(n * 4) + 2 * al(m * 4) + al(chunks * 4);
}

template <typename K>
hipError_t radix_sort_pairs(const K* keys_in, K* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                            int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t tiles = (n + kTileKeys - 1) / kTileKeys;
  const int64_t m = tiles * 256;
  char* w = static_cast<char*>(ws);
  K* ktmp = reinterpret_cast<K*>(w);
  w += al(n * sizeof(K));
  uint32_t* vtmp = reinterpret_cast<uint32_t*>(w);
  w += al(n * 4);
  uint32_t* counts = reinterpret_cast<uint32_t*>(w);
  w += al(m * 4);
  uint32_t* tile_counts = reinterpret_cast<uint32_t*>(w);
  w += al(m * 4);
  uint32_t* sums = reinterpret_cast<uint32_t*>(w);
  const int passes = (end_bit - begin_bit + 7) / 8;
  const K* ksrc = keys_in;
  const uint32_t* vsrc = vals_in;  // null: the pass numbers the input positions
  for (int p = 0; p < passes; ++p) {
    const int shift = begin_bit + 8 * p;
    // the last pass lands in the outputs
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    K* kdst = to_out ? keys_out : ktmp;
    uint32_t* vdst = to_out ? vals_out : vtmp;
    hipLaunchKernelGGL((digit_hist_kernel<K>), dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0, st,
                       ksrc, n, shift, tiles, counts, tile_counts);
    hipError_t e = exclusive_scan(counts, m, sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((digit_scatter_kernel<K>), dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0, st,
                       ksrc, vsrc, n, shift, tiles, counts, tile_counts, kdst, vdst);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipSuccess;
}

template hipError_t radix_sort_pairs<uint32_t>(const uint32_t*, uint32_t*, const uint32_t*, uint32_t*, int64_t,
                                               int, int, void*, hipStream_t);
template hipError_t radix_sort_pairs<uint64_t>(const uint64_t*, uint64_t*, const uint32_t*, uint32_t*, int64_t,
                                               int, int, void*, hipStream_t);

}  // namespace aknn
