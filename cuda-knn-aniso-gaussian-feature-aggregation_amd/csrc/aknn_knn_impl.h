// Brute-force anisotropic top-K for gfx950.
//
// Replaces the neighbour scan of knn_aggregate_aniso_forward_cuda_kernel
// (csrc/forward.cu:83-97) and its MinK selection (csrc/mink.cuh:38-121).
//
// Layout and mapping:
//   * Point rows are packed once per call into 16-value records
//     {p0 p1 p2, R00..R22, s0 s1 s2, 0} (64 B in FP32).  The scan walks the records
//     with a wave-uniform index, so they are fetched with scalar loads and feed the
//     VALU as SGPR operands: no LDS traffic and no per-lane address math.
//   * Each lane owns QPT queries and keeps, per query, a sorted register list of the
//     KMAX best (key, index) pairs.  A candidate is inserted only when its key is below
//     the list's last key; that test is the per-pair cost besides the 26-op key.
//   * SPLIT waves of a workgroup share the same queries and scan disjoint contiguous
//     ranges of the cloud's points; their lists are merged through LDS in a fixed
//     tree order.  Small query counts use SPLIT > 1 to fill the 256 CUs.
//   * Order: ascending key, equal keys by ascending point index (the (key, index)
//     total order).  Because each wave scans its range in ascending index order, a
//     later candidate with an equal key never displaces an earlier one, so the scan
//     compares keys only; the fill and merge steps compare (key, index).
#pragma once

#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

constexpr int kSentinel = kNoPoint;

template <typename T>
__device__ __forceinline__ T pos_inf();
template <>
__device__ __forceinline__ float pos_inf<float>() { return __builtin_inff(); }
template <>
__device__ __forceinline__ double pos_inf<double>() { return __builtin_inf(); }

// ------------------------------------------------------------------ sorted lists
// Insert (c, ci) into the ascending list (k, v) of length N, dropping the last entry.
// lt[i] says whether the candidate goes before slot i; it is computed for every slot
// against the candidate itself (not against a carried value), so it is monotone and
// equal keys already in the list keep their relative order.  The shift then moves
// one carried value down the list: straight-line compares and selects, no branches.
// LEX selects the full (key, index) comparison; otherwise an equal key already in the
// list stays ahead of the candidate.
template <int N, bool LEX, typename T>
__device__ __forceinline__ void list_insert(T (&k)[N], int (&v)[N], T c, int ci) {
  bool lt[N];
#pragma unroll
  for (int i = 0; i < N; ++i) lt[i] = LEX ? ((c < k[i]) | ((c == k[i]) & (ci < v[i]))) : (c < k[i]);
  T carry = c;
  int cv = ci;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const T kk = k[i];
    const int vv = v[i];
    k[i] = lt[i] ? carry : kk;
    v[i] = lt[i] ? cv : vv;
    carry = lt[i] ? kk : carry;
    cv = lt[i] ? vv : cv;
  }
}

// ------------------------------------------------------------------ top-K kernel
// One workgroup = SPLIT waves = one tile of 64*QPT queries of one cloud.
// grid.x = query tiles (dense: per cloud; ragged: from the tile table), grid.y = clouds
// (dense) or 1 (ragged).
template <typename T, int KMAX, int QPT, int SPLIT, int KEY>
__global__ __launch_bounds__(64 * SPLIT) void knn_topk_kernel(
    const T* __restrict__ q, const T* __restrict__ rec, Clouds cl, const int64_t* __restrict__ tiles,
    int K, T* __restrict__ dists, int64_t* __restrict__ idxs) {
  // wave index made visibly wave-uniform, so the point walk below stays scalar
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;

  int64_t b, q0;
  if (tiles) {
    // ragged: tiles[b] = first tile of cloud b (B+1 entries); find the last cloud
    // whose first tile is <= this block (empty clouds share their successor's start)
    const int64_t t = blockIdx.x;
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (tiles[mid] <= t) lo = mid; else hi = mid - 1;
    }
    b = lo;
    q0 = cl.q_begin(b) + (t - tiles[b]) * (64 * QPT);
  } else {
    b = blockIdx.y;
    q0 = cl.q_begin(b) + static_cast<int64_t>(blockIdx.x) * (64 * QPT);
  }
  const int64_t qend = cl.q_end(b);
  const int64_t pb = cl.p_begin(b);
  const int np = static_cast<int>(cl.p_end(b) - pb);

  T qx[QPT], qy[QPT], qz[QPT];
#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int64_t row = q0 + lane + 64 * i;
    if (row < qend) {
      qx[i] = q[row * 3 + 0];
      qy[i] = q[row * 3 + 1];
      qz[i] = q[row * 3 + 2];
    } else {  // NaN keys never enter a list
      qx[i] = qy[i] = qz[i] = __builtin_nan("");
    }
  }

  T key[QPT][KMAX];
  int id[QPT][KMAX];
#pragma unroll
  for (int i = 0; i < QPT; ++i)
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      key[i][s] = pos_inf<T>();
      id[i][s] = kSentinel;
    }

  // this wave's contiguous point range
  const int chunk = (np + SPLIT - 1) / SPLIT;
  const int lo = min(np, wave * chunk);
  const int hi = min(np, lo + chunk);
  const T* __restrict__ base = rec + pb * 16;

  // fill: the first KMAX points enter unconditionally (ordered by (key, index))
  const int fill_end = min(hi, lo + KMAX);
  for (int n = lo; n < fill_end; ++n) {
    const T* r = base + static_cast<int64_t>(n) * 16;
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const T c = pair_key<T, KEY>(qx[i], qy[i], qz[i], r[0], r[1], r[2], r[3], r[4], r[5], r[6],
                                   r[7], r[8], r[9], r[10], r[11], r[12], r[13], r[14]);
      list_insert<KMAX, true>(key[i], id[i], c, n);
    }
  }
  // scan
  for (int n = fill_end; n < hi; ++n) {
    const T* r = base + static_cast<int64_t>(n) * 16;
    const T p0 = r[0], p1 = r[1], p2 = r[2];
    const T r00 = r[3], r01 = r[4], r02 = r[5], r10 = r[6], r11 = r[7], r12 = r[8];
    const T r20 = r[9], r21 = r[10], r22 = r[11], s0 = r[12], s1 = r[13], s2 = r[14];
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const T c = pair_key<T, KEY>(qx[i], qy[i], qz[i], p0, p1, p2, r00, r01, r02, r10, r11, r12,
                                   r20, r21, r22, s0, s1, s2);
      if (c < key[i][KMAX - 1]) list_insert<KMAX, false>(key[i], id[i], c, n);
    }
  }

  if constexpr (SPLIT > 1) {
    // tree merge: at step h, waves w with w % (2h) == h hand their lists to wave w - h
    __shared__ T skey[SPLIT / 2][QPT][KMAX][64];
    __shared__ int sid[SPLIT / 2][QPT][KMAX][64];
#pragma unroll 1
    for (int h = 1; h < SPLIT; h <<= 1) {
      const int slot = wave / (2 * h);
      if ((wave % (2 * h)) == h) {
#pragma unroll
        for (int i = 0; i < QPT; ++i)
#pragma unroll
          for (int s = 0; s < KMAX; ++s) {
            skey[slot][i][s][lane] = key[i][s];
            sid[slot][i][s][lane] = id[i][s];
          }
      }
      __syncthreads();
      if ((wave % (2 * h)) == 0) {
#pragma unroll
        for (int i = 0; i < QPT; ++i)
#pragma unroll 1
          for (int s = 0; s < KMAX; ++s) {
            const T c = skey[slot][i][s][lane];
            const int ci = sid[slot][i][s][lane];
            const bool better = (c < key[i][KMAX - 1]) |
                                ((c == key[i][KMAX - 1]) & (ci < id[i][KMAX - 1]));
            if (better) list_insert<KMAX, true>(key[i], id[i], c, ci);
          }
      }
      __syncthreads();
    }
    if (wave != 0) return;
  }

#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int64_t row = q0 + lane + 64 * i;
    if (row < qend) {
      int last = 0;
#pragma unroll
      for (int s = 0; s < KMAX; ++s) {
        if (s < K) {
          dists[row * K + s] = key[i][s];
          idxs[row * K + s] = static_cast<int64_t>(id[i][s]);
          last = id[i][s];
        }
      }
      if (last == kNoPoint) fill_unfilled_row<T>(dists + row * K, idxs + row * K, K);
    }
  }
}

// ------------------------------------------------------------------ launch
template <typename T, int KMAX, int QPT, int SPLIT, int KEY>
static hipError_t launch_topk(const KnnLaunch& L, hipStream_t st) {
  const dim3 block(64 * SPLIT);
  dim3 grid;
  if (L.tiles) grid = dim3(static_cast<unsigned>(L.ntiles), 1);
  else grid = dim3(static_cast<unsigned>((L.cl.Q + 64 * QPT - 1) / (64 * QPT)), static_cast<unsigned>(L.cl.B));
  hipLaunchKernelGGL((knn_topk_kernel<T, KMAX, QPT, SPLIT, KEY>), grid, block, 0, st,
                     static_cast<const T*>(L.q), static_cast<const T*>(L.rec), L.cl, L.tiles,
                     static_cast<int>(L.K), static_cast<T*>(L.dists), L.idxs);
  return hipGetLastError();
}

// LDS of the merge tree: SPLIT/2 staged lists of QPT x KMAX (key, index) per lane.
template <typename T, int KMAX, int QPT, int SPLIT>
constexpr int merge_lds_bytes() {
  return SPLIT > 1 ? (SPLIT / 2) * QPT * KMAX * 64 * static_cast<int>(sizeof(T) + sizeof(int)) : 0;
}
constexpr int kMergeLdsLimit = 64 * 1024;

template <typename T, int KMAX, int QPT, int KEY>
static hipError_t dispatch_split(const KnnLaunch& L, hipStream_t st) {
  // a split whose merge buffers do not fit falls back to the largest one that does
  if (L.split >= 8) {
    if constexpr (merge_lds_bytes<T, KMAX, QPT, 8>() <= kMergeLdsLimit) return launch_topk<T, KMAX, QPT, 8, KEY>(L, st);
  }
  if (L.split >= 4) {
    if constexpr (merge_lds_bytes<T, KMAX, QPT, 4>() <= kMergeLdsLimit) return launch_topk<T, KMAX, QPT, 4, KEY>(L, st);
  }
  if (L.split >= 2) {
    if constexpr (merge_lds_bytes<T, KMAX, QPT, 2>() <= kMergeLdsLimit) return launch_topk<T, KMAX, QPT, 2, KEY>(L, st);
  }
  return launch_topk<T, KMAX, QPT, 1, KEY>(L, st);
}

// Several queries per lane are compiled for the FP32 anisotropic key only (the
// production path); other variants run one query per lane.
template <typename T, int KMAX, int KEY>
static hipError_t dispatch_qpt(const KnnLaunch& L, hipStream_t st) {
  if constexpr (sizeof(T) == 4 && KEY == AKNN_KEY_ANISO && KMAX <= 16) {
    if (L.qpt == 2) return dispatch_split<T, KMAX, 2, KEY>(L, st);
    if (L.qpt >= 4) return dispatch_split<T, KMAX, 4, KEY>(L, st);
  }
  return dispatch_split<T, KMAX, 1, KEY>(L, st);
}

// The list length KMAX is K rounded up to a compiled size; the kernel keeps the KMAX
// best candidates (a superset of the K best) and stores the first K.
template <typename T, int KEY>
hipError_t launch_knn_typed(const KnnLaunch& L, hipStream_t st) {
  const int64_t K = L.K;
  if (K <= 1) return dispatch_qpt<T, 1, KEY>(L, st);
  if (K <= 4) return dispatch_qpt<T, 4, KEY>(L, st);
  if (K <= 8) return dispatch_qpt<T, 8, KEY>(L, st);
  if (K <= 16) return dispatch_qpt<T, 16, KEY>(L, st);
  if (K <= 32) return dispatch_qpt<T, 32, KEY>(L, st);
  return dispatch_qpt<T, 64, KEY>(L, st);
}

}  // namespace aknn
