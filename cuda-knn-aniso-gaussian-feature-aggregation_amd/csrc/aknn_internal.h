// Host-side launch descriptors shared by the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aknn_common.h"

namespace aknn {

struct KnnLaunch {
  int32_t dtype, key;
  Clouds cl;
  const int64_t* tiles;  // ragged tile table (device), or nullptr for dense grids
  int64_t ntiles;
  int64_t K;
  int qpt, split;
  const void* q;
  const void* rec;  // packed point records [rows][16]
  void* dists;
  int64_t* idxs;
};

// Point tiles of the tiled search and the per-query walk: records per tile, and tiles
// per super-tile (default; the tuner takes 8, 16 or 32).  Queries go in tiles of 64
// (one per lane of a wave).  Sub-tiles of the tiled search: 8 records by default (the
// tuner takes 4, 8 or 16).  The smallest sizes set the header arrays' sizes.
constexpr int kTile = 64;
constexpr int kSuper = 16;
constexpr int kSuperMin = 8;
constexpr int kSubRecDefault = 8;
constexpr int kMuClassesDefault = 1;  // record classes by bound within a tile
constexpr int kSubRecMin = 4;
constexpr int kQueryTile = 64;

// Tile-pruned top-K (same results as KnnLaunch's brute-force scan).
struct PrunedLaunch {
  int32_t dtype, key;
  Clouds cl;
  int64_t qrows, prows;   // query and point rows of the batch
  int64_t ntiles;         // point tiles (sum over clouds of ceil(P_b / 64))
  int64_t nsuper;         // super-tiles (sum over clouds of ceil(tiles_b / kSuper))
  int64_t nqtiles;        // query tiles (ragged: sum of ceil(Q_b / 64))
  int64_t max_points;     // largest cloud (sizes the box reduction grid)
  int32_t ways;           // blocks sharing one query tile's walk (power of two <= 64)
  int32_t ways_max;       // largest split the workspace is sized for
  unsigned long long* stats;  // optional device counters (super tests, tile tests, scans, waves)
  int64_t K;
  const void *q, *p, *sigma, *R;
  void* dists;
  int64_t* idxs;
  void* ws;
  size_t ws_bytes;
  const void* index;      // prebuilt point index (pruned_index_bytes), or null: built per call
  int32_t wave;           // per-query walk (one wave per query) instead of query tiles
  int32_t super_tiles;    // tiles per super-tile (8, 16 or 32)
  int32_t sub_rec;        // records per sub-tile of the tiled search (4, 8 or 16)
  int32_t waves;          // tiled search occupancy target (waves per SIMD; 0 = the list's default)
  int32_t mu_classes;     // record classes by bound within a tile (1 = Morton order)
  int32_t persistent;     // tiled search: resident blocks over a work list of query tiles
  int64_t nsuper_alloc;   // super-tiles at the smallest super-tile size (sizes the headers)
};

// Gather-side (query rows) work: forward aggregation, first- and second-order dq.
struct QueryLaunch {
  int32_t dtype, key;
  Clouds cl;
  int64_t rows;  // query rows in the batch
  int64_t F, K;
  const void *q, *p, *f, *sigma, *R;
  const int64_t* idxs;
  const void* dists;       // forward: keys from the top-K (nullptr: recompute from idxs)
  void* dists_out;         // forward with recomputed keys: where to store them
  const void* grad_f_out;  // backward
  const void* grad_grad_q; // second order (may be null in the full second order)
  void *f_out, *w_out;     // forward
  void* dq;                // backward (first or second order)
  void* d_grad_f_out;      // second order
  const int32_t* order;    // optional processing order of the query rows (spatial)
  const void* rec;         // optional packed point records [rows][16]
  // (kept last: the argument block of the other modes is unchanged)
  const void *gg_p, *gg_sigma, *gg_R;  // full second order: upstream of dp, dsigma, dR
  bool accumulate;         // forward: add into f_out instead of overwriting it
  bool g_bcast;            // grad_f_out holds one row per cloud ([B, F]) instead of per query
  const void* dots;        // dq from point dot products: a[n] = grad_f_out[cloud(n)] . f[n]
  const void* gw;          // optional grad_w_out [rows][K]: added to each slot's a
  void* dgw;               // optional second-order output d grad_w_out [rows][K]
};

// Point-side work over the CSR transpose of idxs.
struct PointLaunch {
  int32_t dtype, key;
  int64_t rows;  // point rows in the batch
  int64_t F, K;
  const int32_t* csr_off;
  const int32_t* csr_ent;
  const void *q, *p, *f, *sigma, *R;
  const void* grad_f_out;
  const void* grad_grad_q;  // second order (null in the full second order when absent)
  void *dp, *df, *dsigma, *dR;
  const int32_t* order;     // optional processing order of the point rows (spatial)
  void* dq_contrib;         // optional [query rows * K * 3]: each entry's share of dq
  // (kept last: the argument block of the first- and second-order kernels is unchanged)
  const void *gg_p, *gg_f, *gg_sigma, *gg_R;  // full second order: upstream of dp, df, dsigma, dR
  bool second;              // second order (grad_grad_q and the gg_* may each be null)
  bool g_bcast;             // grad_f_out holds one row per cloud ([B, F]) instead of per query
  Clouds cl;                // cloud of a point row (broadcast upstream)
  const void* gw;           // optional grad_w_out [query rows * K]: added to each entry's a
  // CSR in a spatial query numbering (entries pos*K + k, qorder[pos] = query row): the
  // query coordinates and grad_grad_q rows are read from copies in that order
  const int32_t* qorder;
  const void* q_sorted;
  const void* hh_sorted;
};

template <typename T>
hipError_t launch_pack_points(const void* p, const void* sigma, const void* R, void* rec,
                              int64_t rows, hipStream_t st);
hipError_t launch_knn(const KnnLaunch& L, hipStream_t st);
hipError_t launch_aggregate(const QueryLaunch& L, hipStream_t st);
// isotropic forward: L.sigma = one scale per point row, L.w_out = [rows] weight sums
hipError_t launch_aggregate_iso(const QueryLaunch& L, hipStream_t st);
hipError_t launch_lookup_gather(const void* d1, const int64_t* i1, const int64_t* table, int64_t table_k,
                                int64_t rows, int64_t K, int32_t dtype, int64_t* idxs, hipStream_t st);
hipError_t launch_query_grad(const QueryLaunch& L, bool second_order, hipStream_t st);
// full second order: d grad_f_out = sum_k w Phi f[n] (no dq)
hipError_t launch_query_grad_full(const QueryLaunch& L, hipStream_t st);
hipError_t launch_point_grad(const PointLaunch& L, hipStream_t st);
// broadcast upstream: a[n] = grad_f_out[cloud(n)] . f[n] for every point row (L.g_bcast),
// in the association of the point kernels
hipError_t launch_point_dots(const PointLaunch& L, void* a, hipStream_t st);
// dq (first or second order) per query from the point dot products L.dots, summed in slot
// order as the per-entry shares are
hipError_t launch_dq_dots(const QueryLaunch& L, bool second_order, hipStream_t st);

// The forward's cache (query order, point order, point records; each may be null) in one
// launch; a null order is written as the identity (aknn_knn.hip).
template <typename T>
hipError_t launch_cache_fill(const uint32_t* qord, int32_t* qout, int64_t qrows, const uint32_t* pord,
                             int32_t* pout, int64_t prows, const void* p, const void* sigma, const void* R,
                             void* rec, hipStream_t st);
hipError_t pruned_workspace(const PrunedLaunch& L, size_t* bytes);
size_t pruned_index_bytes(const PrunedLaunch& L);
size_t pruned_index_build_ws(const PrunedLaunch& L);
hipError_t launch_pruned_index_build(const PrunedLaunch& L, hipStream_t st);  // L.index = destination
hipError_t launch_knn_pruned_prep(const PrunedLaunch& L, hipStream_t st);
hipError_t launch_knn_pruned_search(const PrunedLaunch& L, hipStream_t st);
// spatial orders of the query and point rows (valid after the prep, inside the workspace)
const uint32_t* pruned_query_order(const PrunedLaunch& L);
const uint32_t* pruned_point_order(const PrunedLaunch& L);

hipError_t launch_contrib_reduce(const void* contrib, int64_t rows, int64_t K, void* dq, int32_t dtype,
                                 hipStream_t st);

// Stable LSD radix sort (aknn_sort.hip): vals_in NULL sorts the positions 0..n-1.
size_t radix_sort_workspace(int64_t n, size_t key_bytes);
template <typename K>
hipError_t radix_sort_pairs(const K* keys_in, K* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                            int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t st,
                            const int32_t* vrank = nullptr, uint32_t vk = 1);

// CSR transpose by two passes through packed words (aknn_sort.hip), for point keys of 9
// to 16 bits whose entry numbers fit beside the high byte: off[0..rows] and the
// entries' values (entry numbers, or vrank[q]*K + k), as the pair sort would give.
bool csr_packed_fits(int64_t n, int64_t K, int key_bits);
size_t csr_packed_workspace(int64_t n);
hipError_t csr_sort_packed(const int64_t* idxs, const Clouds& cl, int64_t K, int64_t n, int64_t rows,
                           const int32_t* vrank, int32_t* off, uint32_t* ent, void* ws, hipStream_t st);

// CSR build
hipError_t csr_temp_bytes(int64_t entries, int64_t point_rows, int64_t query_rows, size_t* bytes);
// qorder (optional): entries number (position in qorder) * K + k instead of q * K + k
// (each row's entries stay in ascending q * K + k order)
hipError_t launch_build_csr(const int64_t* idxs, Clouds cl, int64_t query_rows, int64_t K,
                            int64_t point_rows, int32_t* csr_off, int32_t* csr_ent, void* ws,
                            size_t ws_bytes, hipStream_t st, const int32_t* qorder = nullptr);
// out[i][0..2] = src[order[i]][0..2] (query-side rows in spatial order)
hipError_t launch_gather3(const void* src, const int32_t* order, void* out, int64_t rows, int32_t dtype,
                          hipStream_t st);

}  // namespace aknn
