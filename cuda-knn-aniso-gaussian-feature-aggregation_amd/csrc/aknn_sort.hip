// Stable LSD radix sort of (key, uint32 value) pairs for gfx950, 8 bits per pass.
//
// Used for the Morton orders of points and queries and for the CSR transpose of the
// neighbour table.  Everything is stream-ordered kernels over caller-provided
// workspace: no allocation, no memset, no host synchronisation, so the whole sort can
// be captured in a graph.  Stable and deterministic: equal keys keep input order.
//
// Per pass over digit (key >> shift) & 255, with tiles of tile_keys<K>() keys per block:
//   1. digit_hist: per-tile digit counts (LDS), stored digit-major: counts[d * tiles + t]
//   2. digit_scan: one block per digit scans that digit's tile counts (in place) and
//      records the digit's total
//   3. digit_scatter: the digit bases (a scan of the 256 totals, per block), local digit
//      starts from the tile's own counts (kept by step 1);
//      each wave ranks its own contiguous run of the tile's keys in order; a key's rank
//      among equal digits is (earlier waves' counts) + (earlier chunks of its wave) +
//      (earlier lanes of its chunk, from 8 ballots over the digit bits).  Keys are
//      placed in LDS at their tile-local rank, then written out run by run.
#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

constexpr int kSortThreads = 256;
// Keys per thread: 16 for 32-bit keys (4096-key tiles), 8 for 64-bit keys (2048-key
// tiles).  The 64-bit sorts are the Morton orders, one key per query or point: a 1M-query
// sort gets 512 blocks instead of 256, enough to give every CU more than one.
template <typename K>
constexpr int sort_items() { return sizeof(K) == 8 ? 8 : 16; }
template <typename K>
constexpr int tile_keys() { return kSortThreads * sort_items<K>(); }

template <typename K>
__global__ __launch_bounds__(kSortThreads) void digit_hist_kernel(const K* __restrict__ keys, int64_t n,
                                                                  int shift, int64_t tiles,
                                                                  uint32_t* __restrict__ counts,
                                                                  uint32_t* __restrict__ tile_counts) {
  constexpr int kSortItems = sort_items<K>();
  constexpr int kTileKeys = tile_keys<K>();
  constexpr int kVec = 16 / static_cast<int>(sizeof(K));  // keys per 16-byte load
  constexpr int kLoads = kSortItems / kVec;
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTileKeys;
  // integer counts: the order of the additions does not matter
  if (base + kTileKeys <= n && (reinterpret_cast<uintptr_t>(keys) % 16) == 0) {
    // whole tile, 16-byte loads (coalesced: thread t takes vectors t, t + 256, ...), all
    // issued before the first count
    const uint4* kv = reinterpret_cast<const uint4*>(keys + base);
    uint4 r[kLoads];
#pragma unroll
    for (int i = 0; i < kLoads; ++i) r[i] = kv[i * kSortThreads + threadIdx.x];
#pragma unroll
    for (int i = 0; i < kLoads; ++i) {
      if constexpr (sizeof(K) == 4) {
        atomicAdd(&h[(r[i].x >> shift) & 255u], 1u);
        atomicAdd(&h[(r[i].y >> shift) & 255u], 1u);
        atomicAdd(&h[(r[i].z >> shift) & 255u], 1u);
        atomicAdd(&h[(r[i].w >> shift) & 255u], 1u);
      } else {
        const uint64_t k0 = static_cast<uint64_t>(r[i].x) | (static_cast<uint64_t>(r[i].y) << 32);
        const uint64_t k1 = static_cast<uint64_t>(r[i].z) | (static_cast<uint64_t>(r[i].w) << 32);
        atomicAdd(&h[(k0 >> shift) & 255u], 1u);
        atomicAdd(&h[(k1 >> shift) & 255u], 1u);
      }
    }
  } else {
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
      const int64_t j = base + i * kSortThreads + threadIdx.x;
      if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);
    }
  }
  __syncthreads();
  counts[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];       // scanned next
  tile_counts[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x] = h[threadIdx.x];  // kept per tile
}

// ---- per-digit scan of the tile counts
// Exclusive scan of 256 values over the block (256 threads); returns the total.
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

// One block per digit d: counts[d * tiles + t] becomes the exclusive prefix over t, in
// place, and totals[d] the digit's count.
constexpr int kScanItems = 8;  // consecutive counts per thread and round
__global__ __launch_bounds__(256) void digit_scan_kernel(uint32_t* __restrict__ counts, int64_t tiles,
                                                         uint32_t* __restrict__ totals) {
  __shared__ uint32_t s[256];
  uint32_t* __restrict__ row = counts + static_cast<int64_t>(blockIdx.x) * tiles;
  uint32_t carry = 0;
  for (int64_t base = 0; base < tiles; base += 256 * kScanItems) {
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const int64_t j = base + static_cast<int64_t>(threadIdx.x) * kScanItems + i;
      v[i] = j < tiles ? row[j] : 0u;
      local += v[i];
    }
    uint32_t ex;
    const uint32_t total = block_exscan256(s, local, &ex);
    uint32_t run = carry + ex;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const int64_t j = base + static_cast<int64_t>(threadIdx.x) * kScanItems + i;
      if (j < tiles) row[j] = run;
      run += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// ---- stable scatter through LDS
// The tile is ranked stably into LDS (digit-major, input order within a digit), then
// written out in LDS order: consecutive threads write consecutive addresses of one
// digit's run, so the global stores coalesce.
//
// kMode selects what goes out (32-bit keys only for modes 1 and 2):
//   0  key and value pairs;
//   1  one packed word per entry, ((key >> 8) << vbits) | value, to kout: the first of
//      two passes over keys of at most 16 bits whose values fit below bit vbits;
//   2  the value bits (word & (2^vbits - 1)) of packed words to vout: the second pass,
//      whose digit is the word's top byte (shift = vbits).
template <typename K, int kMode>
__global__ __launch_bounds__(kSortThreads) void digit_scatter_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift, int64_t tiles,
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ totals, const uint32_t* __restrict__ tile_counts,
    K* __restrict__ kout, uint32_t* __restrict__ vout, const int32_t* __restrict__ vrank, uint32_t vk, int vbits) {
  static_assert(kMode == 0 || sizeof(K) == 4, "packed passes take 32-bit words");
  constexpr int kSortItems = sort_items<K>();
  constexpr int kTileKeys = tile_keys<K>();
  constexpr bool kVals = kMode != 2;  // mode 2 carries no separate value
  // mode 1 places the packed word, and beside it the digit it no longer holds
  __shared__ K skey[kTileKeys];
  __shared__ uint32_t sval[kMode == 0 ? kTileKeys : 1];
  __shared__ uint8_t sdig[kMode == 1 ? kTileKeys : 1];
  __shared__ uint32_t loc[256];        // local start of each digit in the tile
  __shared__ uint32_t gpos[256];       // global start of each digit's run of this tile
  __shared__ uint32_t wcnt[4][256];    // per-wave digit counts, then each wave's digit starts
  __shared__ uint32_t scan_tmp[256];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // Each XCD takes one contiguous range of tiles: the runs that adjacent tiles write for
  // one digit are adjacent in the output, and the line at their junction is then
  // completed in one L2.
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const int64_t base = tile * kTileKeys;
  const int tile_n = static_cast<int>(min<int64_t>(kTileKeys, n - base));

  // local digit starts from the tile's histogram (digit_hist), global starts from the scan
  uint32_t ex;
  block_exscan256(scan_tmp, tile_counts[tile * 256 + threadIdx.x], &ex);
  loc[threadIdx.x] = ex;
  uint32_t dbase;  // this digit's first output position: the counts of the lower digits
  block_exscan256(scan_tmp, totals[threadIdx.x], &dbase);
  gpos[threadIdx.x] = dbase + start[static_cast<int64_t>(threadIdx.x) * tiles + tile];
#pragma unroll
  for (int w = 0; w < 4; ++w) wcnt[w][threadIdx.x] = 0;

  // Each wave ranks its own contiguous run of kSortItems * 64 keys of the tile, in input
  // order: a key's rank within its wave's run for its digit is the wave's count of that
  // digit so far (in LDS; the wave's LDS operations complete in program order) plus the
  // earlier lanes of its chunk with the same digit (8 ballots over the digit bits).  No
  // block barrier inside the run.  All the run's keys and values are loaded up front.
  const int wbase = wave * (kSortItems * 64);
  const FastDiv vdiv(vk ? vk : 1u);
  K kreg[kSortItems];
  uint32_t vreg[kSortItems], lrank[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int jj = wbase + i * 64 + lane;
    const bool ok = jj < tile_n;
    kreg[i] = ok ? kin[base + jj] : K(0);
    vreg[i] = 0u;
    if constexpr (kVals) {
      vreg[i] = ok ? (vin ? vin[base + jj] : static_cast<uint32_t>(base + jj)) : 0u;
      if (!vin && vrank && ok) {  // position n = r*vk + k becomes vrank[r]*vk + k
        const uint32_t n = vreg[i];
        const uint32_t r = vdiv.div(n);
        vreg[i] = static_cast<uint32_t>(vrank[r]) * vk + (n - r * vk);
      }
    }
  }
  __syncthreads();  // loc, gpos and the zeroed counts are visible
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const bool ok = wbase + i * 64 + lane < tile_n;
    const uint32_t d = ok ? static_cast<uint32_t>((kreg[i] >> shift) & 255u) : 256u;
    // lanes with the same digit: start from the lanes holding a key, and for each digit
    // bit keep the lanes whose bit agrees with this lane's (the ballot of the bit, or
    // its complement); bm is the lane's bit as an all-ones or all-zeros word, so each
    // half of the mask takes one three-input bit operation per bit
    const uint64_t okm = __builtin_amdgcn_ballot_w64(ok);
    uint32_t slo = static_cast<uint32_t>(okm), shi = static_cast<uint32_t>(okm >> 32);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t bm = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - b)) >> 31);
      const uint64_t m = __builtin_amdgcn_ballot_w64(bm != 0u);
      slo &= ~(static_cast<uint32_t>(m) ^ bm);
      shi &= ~(static_cast<uint32_t>(m >> 32) ^ bm);
    }
    // rank among the earlier lanes of the chunk with the same digit, and their count
    const uint32_t rank_in = __builtin_amdgcn_mbcnt_hi(shi, __builtin_amdgcn_mbcnt_lo(slo, 0u));
    const uint32_t prev = ok ? wcnt[wave][d] : 0u;
    lrank[i] = prev + rank_in;
    if (ok && rank_in == 0) wcnt[wave][d] = prev + static_cast<uint32_t>(__popc(slo) + __popc(shi));
    __builtin_amdgcn_wave_barrier();  // the next chunk's reads follow this chunk's update
  }
  __syncthreads();
  {  // digit t: each wave's start = the tile's digit start + the earlier waves' counts
    uint32_t acc = loc[threadIdx.x];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t c = wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    if (wbase + i * 64 + lane < tile_n) {
      const uint32_t d = static_cast<uint32_t>((kreg[i] >> shift) & 255u);
      const uint32_t off = wcnt[wave][d] + lrank[i];
      if constexpr (kMode == 1) {
        skey[off] = ((static_cast<uint32_t>(kreg[i]) >> 8) << vbits) | vreg[i];
        sdig[off] = static_cast<uint8_t>(d);
      } else {
        skey[off] = kreg[i];
        if constexpr (kVals) sval[off] = vreg[i];
      }
    }
  }
  __syncthreads();
  // coalesced write-out of the locally sorted tile
  for (int jj = threadIdx.x; jj < tile_n; jj += kSortThreads) {
    const K key = skey[jj];
    const uint32_t d = kMode == 1 ? sdig[jj] : static_cast<uint32_t>((key >> shift) & 255u);
    const uint32_t pos = gpos[d] + (jj - loc[d]);
    if constexpr (kMode == 0) {
      kout[pos] = key;
      vout[pos] = sval[jj];
    } else if constexpr (kMode == 1) {
      kout[pos] = key;
    } else {
      vout[pos] = static_cast<uint32_t>(key) & ((1u << vbits) - 1u);
    }
  }
}

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

size_t radix_sort_workspace(int64_t n, size_t key_bytes) {
  const int64_t tk = key_bytes == 8 ? tile_keys<uint64_t>() : tile_keys<uint32_t>();
  const int64_t tiles = (n + tk - 1) / tk;
  const int64_t m = tiles * 256;
  return al(n * key_bytes) + al(n * 4) + 2 * al(m * 4) + al(256 * 4);
}

template <typename K>
hipError_t radix_sort_pairs(const K* keys_in, K* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                            int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t st, const int32_t* vrank,
                            uint32_t vk) {
  if (n <= 0) return hipSuccess;
  constexpr int64_t kTileKeys = tile_keys<K>();
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
  uint32_t* totals = reinterpret_cast<uint32_t*>(w);  // [256] per-digit counts
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
    hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(256), 0, st, counts, tiles, totals);
    hipLaunchKernelGGL((digit_scatter_kernel<K, 0>), dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0, st,
                       ksrc, vsrc, n, shift, tiles, counts, totals, tile_counts, kdst, vdst, vrank, vk, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipSuccess;
}

// ---- CSR transpose of a neighbour table with point keys of 9 to 16 bits
// The keys are the global point rows of the entries e = q*K + k; the sort's values the
// entry numbers (or, with vrank, vrank[q]*K + k).  Two 8-bit passes, as above, with one
// 32-bit word per entry in between: the first pass writes (key >> 8) << vbits | value,
// the second takes its digit from the top byte and writes the values alone.  The row
// starts come from the two passes' counts, without a search over sorted keys.

// The first pass's histogram, with the keys made on the way: one 4096-entry tile per
// block, thread t taking the 4-entry groups t, t + 256, ... (K divisible by 4, so a
// group is one query row with one cloud lookup; idxs 16-byte aligned; n below 2^31).
// The keys are written out and the counts stored as digit_hist stores them.
__global__ __launch_bounds__(kSortThreads) void csr_keys_hist_kernel(const int64_t* __restrict__ idxs, Clouds cl,
                                                                     uint32_t K, uint32_t n, uint32_t tiles,
                                                                     uint32_t* __restrict__ keys,
                                                                     uint32_t* __restrict__ counts,
                                                                     uint32_t* __restrict__ tile_counts) {
  constexpr int kTileKeys = tile_keys<uint32_t>();
  constexpr int kGroups = kTileKeys / 4 / kSortThreads;
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const FastDiv kdiv(K);
  const uint32_t base = blockIdx.x * kTileKeys;
  uint4 kv[kGroups];
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    const uint32_t e = base + 4u * (g * kSortThreads + threadIdx.x);
    kv[g] = make_uint4(0u, 0u, 0u, 0u);
    if (e < n) {
      const uint32_t pb = static_cast<uint32_t>(cl.p_begin(cl.cloud_of_query(kdiv.div(e))));
      const longlong2 a = *reinterpret_cast<const longlong2*>(idxs + e);
      const longlong2 c = *reinterpret_cast<const longlong2*>(idxs + e + 2);
      kv[g] = make_uint4(pb + static_cast<uint32_t>(a.x), pb + static_cast<uint32_t>(a.y),
                         pb + static_cast<uint32_t>(c.x), pb + static_cast<uint32_t>(c.y));
    }
  }
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    const uint32_t e = base + 4u * (g * kSortThreads + threadIdx.x);
    if (e < n) {
      atomicAdd(&h[kv[g].x & 255u], 1u);
      atomicAdd(&h[kv[g].y & 255u], 1u);
      atomicAdd(&h[kv[g].z & 255u], 1u);
      atomicAdd(&h[kv[g].w & 255u], 1u);
      *reinterpret_cast<uint4*>(keys + e) = kv[g];
    }
  }
  __syncthreads();
  counts[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
  tile_counts[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x] = h[threadIdx.x];
}

// off[k] = the number of entries with a key below k, for k = (hi << 8) | lo, block lo,
// thread hi.  After the first pass the words are ordered by the low byte, so the
// entries with a smaller key are those with a high byte below hi (the second pass's
// digit totals) and those with high byte hi among the first X words, X = the number
// of low bytes below lo: the second pass's scanned count for the tile holding word X,
// plus the words with that high byte from the tile's start to X.
__global__ __launch_bounds__(256) void csr_starts_kernel(const uint32_t* __restrict__ packed, int64_t n, int vbits,
                                                         const uint32_t* __restrict__ totals1,
                                                         const uint32_t* __restrict__ start2,
                                                         const uint32_t* __restrict__ totals2, int64_t tiles,
                                                         int64_t rows, int32_t* __restrict__ off) {
  constexpr int kTileKeys = tile_keys<uint32_t>();
  __shared__ uint32_t s[256], h[256];
  const int lo = blockIdx.x, hi = threadIdx.x;
  uint32_t ignored;
  const uint32_t x = block_exscan256(s, hi < lo ? totals1[hi] : 0u, &ignored);
  const int64_t tile = x / kTileKeys;
  h[hi] = 0;
  __syncthreads();
  for (int64_t j = tile * kTileKeys + hi; j < x; j += 256) atomicAdd(&h[packed[j] >> vbits], 1u);
  __syncthreads();
  uint32_t below;
  block_exscan256(s, totals2[hi], &below);
  const uint32_t c = (tile < tiles ? start2[static_cast<int64_t>(hi) * tiles + tile] : totals2[hi]) + h[hi];
  const int64_t k = (static_cast<int64_t>(hi) << 8) | lo;
  if (k < rows) off[k] = static_cast<int32_t>(below + c);
  if (lo == 0 && hi == 0) off[rows] = static_cast<int32_t>(n);
}

static int bits_for(int64_t count) {  // smallest b with 2^b >= count, at least 1
  int b = 1;
  while ((int64_t(1) << b) < count) ++b;
  return b;
}

bool csr_packed_fits(int64_t n, int64_t K, int key_bits) {
  return n > 0 && n < (int64_t(1) << 31) && K % 4 == 0 && key_bits > 8 && key_bits <= 16 &&
         (key_bits - 8) + bits_for(n) <= 32;
}

size_t csr_packed_workspace(int64_t n) {
  const int64_t tiles = (n + tile_keys<uint32_t>() - 1) / tile_keys<uint32_t>();
  return 2 * al(n * 4) + 2 * al(tiles * 256 * 4) + 2 * al(256 * 4);
}

hipError_t csr_sort_packed(const int64_t* idxs, const Clouds& cl, int64_t K, int64_t n, int64_t rows,
                           const int32_t* vrank, int32_t* off, uint32_t* ent, void* ws, hipStream_t st) {
  constexpr int64_t kTileKeys = tile_keys<uint32_t>();
  const int64_t tiles = (n + kTileKeys - 1) / kTileKeys;
  const int64_t m = tiles * 256;
  const int vbits = bits_for(n);
  char* w = static_cast<char*>(ws);
  uint32_t* keys = reinterpret_cast<uint32_t*>(w);
  w += al(n * 4);
  uint32_t* packed = reinterpret_cast<uint32_t*>(w);
  w += al(n * 4);
  uint32_t* counts = reinterpret_cast<uint32_t*>(w);
  w += al(m * 4);
  uint32_t* tile_counts = reinterpret_cast<uint32_t*>(w);
  w += al(m * 4);
  uint32_t* totals1 = reinterpret_cast<uint32_t*>(w);
  w += al(256 * 4);
  uint32_t* totals2 = reinterpret_cast<uint32_t*>(w);
  const dim3 grid(static_cast<unsigned>(tiles)), block(kSortThreads);
  hipLaunchKernelGGL(csr_keys_hist_kernel, grid, block, 0, st, idxs, cl, static_cast<uint32_t>(K),
                     static_cast<uint32_t>(n), static_cast<uint32_t>(tiles), keys, counts, tile_counts);
  hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(256), 0, st, counts, tiles, totals1);
  hipLaunchKernelGGL((digit_scatter_kernel<uint32_t, 1>), grid, block, 0, st, keys, nullptr, n, 0, tiles, counts,
                     totals1, tile_counts, packed, nullptr, vrank, static_cast<uint32_t>(K), vbits);
  hipLaunchKernelGGL((digit_hist_kernel<uint32_t>), grid, block, 0, st, packed, n, vbits, tiles, counts,
                     tile_counts);
  hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(256), 0, st, counts, tiles, totals2);
  hipLaunchKernelGGL(csr_starts_kernel, dim3(256), dim3(256), 0, st, packed, n, vbits, totals1, counts, totals2,
                     tiles, rows, off);
  hipLaunchKernelGGL((digit_scatter_kernel<uint32_t, 2>), grid, block, 0, st, packed, nullptr, n, vbits, tiles,
                     counts, totals2, tile_counts, nullptr, ent, nullptr, 0u, vbits);
  return hipGetLastError();
}

template hipError_t radix_sort_pairs<uint32_t>(const uint32_t*, uint32_t*, const uint32_t*, uint32_t*, int64_t,
                                               int, int, void*, hipStream_t, const int32_t*, uint32_t);
template hipError_t radix_sort_pairs<uint64_t>(const uint64_t*, uint64_t*, const uint32_t*, uint32_t*, int64_t,
                                               int, int, void*, hipStream_t, const int32_t*, uint32_t);

}  // namespace aknn
