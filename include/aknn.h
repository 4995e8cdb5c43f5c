/*
 * aknn.h -- C ABI of the MI355X (gfx950) library behind knn_aggregate_aniso.
 *
 * The library takes raw device pointers and sizes only.  It allocates nothing:
 * every output and every workspace buffer is owned by the caller (the Python
 * layer allocates them with torch).  All work is enqueued on the caller's HIP
 * stream (passed as `void*`, i.e. a hipStream_t); no call synchronises the
 * device.  Functions return 0 on success and a negative AKNN_E* code on error;
 * aknn_last_error() then returns a message for the calling thread.
 *
 * Tensor layouts (row-major, contiguous), all floating tensors of one dtype:
 *   q      [B, Q, 3]      query coordinates
 *   p      [B, P, 3]      point coordinates
 *   f      [B, P, F]      point features
 *   sigma  [B, P, 3]      per-point Gaussian scales (multiply the squared offset)
 *   R      [B, P, 3, 3]   per-point rotations, R[b][n][row][col]
 *   f_out  [B, Q, F], w_out [B, Q, K], dists [B, Q, K], idxs [B, Q, K] (int64)
 * Ragged batches use the same layouts with the cloud axis folded into the row
 * axis ([sum Q_i, 3], [sum P_i, 3], ...) and device offset arrays of B+1 int64.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   aknn_forward    <- knn_aggregate_aniso_forward        csrc/knn_aggregate.cpp:26-47,
 *                      launcher csrc/forward.cu:286-338
 *   aknn_backward   <- knn_aggregate_aniso_backward       csrc/knn_aggregate.cpp:101-132,
 *                      launcher csrc/backward.cu:331-389
 *   aknn_backward2  <- knn_aggregate_aniso_backward_2nd   csrc/knn_aggregate.cpp:135-172,
 *                      launcher csrc/backward.cu:392-455
 *   aknn_aggregate  <- knn_precomputed_aggregate_aniso_forward  csrc/knn_aggregate.cpp:76-99
 *   aknn_lookup_forward <- knn_lookup_aggregate_aniso_forward  csrc/knn_aggregate.cpp:50-73
 *   aknn_iso_forward    <- knn_aggregate_forward               csrc/knn_aggregate.cpp:5-23
 *   (pybind11 table csrc/knn_aggregate.cpp:178-185)
 */
#ifndef AKNN_H_
#define AKNN_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define AKNN_API __attribute__((visibility("default")))
#else
#define AKNN_API
#endif

#define AKNN_ABI_VERSION 2

/* error codes */
#define AKNN_OK 0
#define AKNN_E_INVALID -1   /* bad argument (shape, K range, null pointer, ...) */
#define AKNN_E_WORKSPACE -2 /* workspace too small */
#define AKNN_E_LAUNCH -3    /* HIP launch / runtime error */
#define AKNN_E_UNSUPPORTED -4 /* beyond this build's limits: K > 64, batch rows >= 2^31, or a feature row
                                wider than one call takes (F > 128 x the widest vector of <= 16 B dividing F:
                                512 in FP32 for F % 4 == 0, 256 for even F, else 128; split wider rows into
                                column ranges, as the Python layer does) */

/* dtype */
#define AKNN_F32 0
#define AKNN_F64 1

/* neighbour key */
#define AKNN_KEY_ANISO 0     /* D = s0 j^2 + s1 k^2 + s2 l^2, (j,k,l) = R (q - p)  (csrc/forward.cu:85-92) */
#define AKNN_KEY_EUCLIDEAN 1 /* D = x^2 + y^2 + z^2                             (csrc/forward.cu:154)   */

/* gradient selection bits for aknn_backward / aknn_backward2 */
#define AKNN_NEED_DQ 1u
#define AKNN_NEED_DP 2u
#define AKNN_NEED_DF 4u
#define AKNN_NEED_DSIGMA 8u
#define AKNN_NEED_DR 16u
#define AKNN_NEED_DGRAD_F_OUT 32u /* aknn_backward2 only */
#define AKNN_SECOND_FULL 64u      /* workspace query for aknn_backward2_full (see there) */
#define AKNN_NEED_POINTS (AKNN_NEED_DP | AKNN_NEED_DF | AKNN_NEED_DSIGMA | AKNN_NEED_DR)
/* Modifier bit for aknn_backward / aknn_backward2 (not aknn_backward2_full): grad_f_out
 * is a broadcast upstream, one row of F values per cloud shared by all of that cloud's
 * query rows, passed as a [B, F] array (the reference materialises such an upstream,
 * e.g. the Eikonal read-out v.expand_as(f_out), with .contiguous() at
 * cuda_knn_aggregate/__init__.py:107).  Results equal those for the materialised
 * [B, Q, F] upstream: the point gradients bit for bit; dq to rounding (it is then summed
 * per query from point dot products a[n] = grad_f_out[cloud(n)] . f[n], see
 * aknn_point_dots, with the bits of the per-entry shares a materialised upstream gets
 * when point gradients are requested too).  The point kernels then form a[n] once per
 * point row instead of gathering one upstream row per neighbour entry. */
#define AKNN_GRAD_BROADCAST 128u

/* Problem description.  Dense batches: q_off = p_off = NULL and B, Q, P give the
 * shape.  Ragged batches: q_off / p_off are DEVICE arrays of B+1 int64 row
 * offsets (q_off[0] = 0, q_off[B] = total query rows), Q and P hold the total row
 * counts.  The host copies of the offsets (q_off_host / p_off_host) are required
 * for ragged launches: they size the launch grid and are validated on the host. */
typedef struct aknn_dims {
  int64_t B, Q, P, F, K;
  int32_t dtype; /* AKNN_F32 | AKNN_F64 */
  int32_t key;   /* AKNN_KEY_* */
  const int64_t* q_off;      /* device, B+1, or NULL (dense) */
  const int64_t* p_off;      /* device, B+1, or NULL (dense) */
  const int64_t* q_off_host; /* host,   B+1, or NULL (dense) */
  const int64_t* p_off_host; /* host,   B+1, or NULL (dense) */
} aknn_dims;

/* Neighbour search algorithms (same results, different cost). */
#define AKNN_SEARCH_AUTO 0
#define AKNN_SEARCH_BRUTE 1  /* scan every point of the cloud                       */
#define AKNN_SEARCH_TILES 2  /* spatially sorted tiles, provably-useless tiles skipped */
#define AKNN_SEARCH_WAVE 3   /* the same tiles walked per query (one wave each): for
                                query sets sparse in their clouds; AUTO picks it there */

/* Tunables of the neighbour search (0 = library default for the shape). */
typedef struct aknn_tune {
  int32_t queries_per_lane; /* brute: 1, 2 or 4 queries held per lane                  */
  int32_t point_split;      /* brute: 1, 2, 4 or 8 waves splitting one tile's points;
                               tiles: blocks sharing one query tile's walk (a power of
                               two, at most the library's own choice for the shape)    */
  int32_t search;           /* AKNN_SEARCH_*                                            */
  int32_t sub_records;      /* tiles: records per sub-tile the search tests and scans (4, 8
                               or 16; FP32 with 9 to 16 neighbours, else 8)             */
  void* stats;              /* tiles: optional device uint64[5] the search adds its counts
                               to (super-tile tests, tile tests, tiles scanned, points
                               scanned, waves); NULL = off                               */
  const void* index;        /* optional point index from aknn_index_build for these
                               points (same dims except the queries): the tiled search
                               then skips its per-call point preparation (and keeps the
                               default tile geometry the index was built with)         */
  int32_t super_tiles;      /* tiles and per-query walk: tiles per super-tile (8, 16, 32) */
  int32_t waves_per_eu;     /* tiles: occupancy target of the search (4 or 8 waves per
                               SIMD; FP32 with 9 to 16 neighbours and 8-record sub-tiles) */
  int32_t mu_classes;       /* tiles and per-query walk: the records of a point tile are
                               stored in 1, 2, 4 or 8 classes by their pruning bound
                               (1 = Morton order), which shapes the sub-tiles       */
  int32_t persistent;       /* tiles: 1 = as many blocks as fit on the device at once,
                               taking (cloud, query tile) work items from per-XCD
                               counters; -1 = one block per work item; 0 = default
                               (1 for ragged batches, -1 for dense ones)           */
} aknn_tune;

/* Point index (no counterpart in the reference; for callers that keep clouds resident
 * and search them many times, like a server).  The tiled search's per-call point
 * preparation -- cloud boxes, point Morton order, tile records and bounds -- built once
 * into a caller-owned device buffer of aknn_index_size bytes (workspace for the build:
 * aknn_index_workspace_size).  Pass it in aknn_tune.index to aknn_forward / aknn_knn
 * with the same points, sigma, R, B, P, offsets, dtype and key mode; results are the
 * same bits as without it.  Rebuild it when the points change. */
AKNN_API size_t aknn_index_size(const aknn_dims* dims);
AKNN_API size_t aknn_index_workspace_size(const aknn_dims* dims);
AKNN_API int aknn_index_build(const aknn_dims* dims, const void* p, const void* sigma, const void* R,
                              void* index, size_t index_bytes, void* workspace, size_t workspace_bytes,
                              void* stream);

/* Optional per-batch cache that aknn_forward fills and the backward passes read (no
 * counterpart in the reference; for callers that keep state between the passes, like
 * the autograd Functions).  Each pointer may be NULL.  Results never depend on it: it
 * only sets the order rows are processed in (spatial orders keep gathered rows in L2)
 * and lets the backward kernels read point data as packed 16-value records.
 * Sizes: query_order int32[query rows]; point_order int32[point rows];
 * point_records dtype[point rows * 16]. */
typedef struct aknn_cache {
  int32_t* query_order;  /* query rows in spatial order */
  int32_t* point_order;  /* point rows in spatial order */
  void* point_records;
  /* dtype[point rows][16]: the dot records of aknn_point_dots, for backward calls with
   * AKNN_GRAD_BROADCAST and the same grad_f_out rows and points (a caller that keeps
   * both fixed, like a server with a fixed read-out, makes them once); NULL = made per
   * call in the workspace.  Not filled by aknn_forward. */
  const void* point_dots;
} aknn_cache;

AKNN_API int aknn_abi_version(void);
AKNN_API const char* aknn_last_error(void);

/* Launch tracing (no counterpart in the reference, which has none).  While enabled,
 * each kernel launch of the library is bracketed by hipEvents on its stream.
 * aknn_trace_collect waits for the recorded events, writes up to max_records
 * NUL-separated launch names into `names` and their durations (ms) into `ms`, returns
 * the count and clears the record.  Not for use under stream capture. */
AKNN_API void aknn_trace_enable(int on);
AKNN_API int aknn_trace_collect(char* names, size_t names_bytes, float* ms, int max_records);

/* Bytes of device workspace aknn_forward / aknn_knn need (enough for either search
 * algorithm).  Returns 0 and sets the error message if the size query fails (for
 * example without a usable device). */
AKNN_API size_t aknn_forward_workspace_size(const aknn_dims* dims);

/* Neighbour search + aggregation: fills f_out, w_out, dists, idxs.  dists holds each
 * neighbour's anisotropic D (w = exp(-D)); with AKNN_KEY_EUCLIDEAN the neighbours are
 * ranked by Euclidean distance but D and w stay anisotropic, like the reference's
 * lookup path (csrc/forward.cu:165-181).
 * Replaces knn_aggregate_aniso_forward (csrc/knn_aggregate.cpp:26-47).  K must be in
 * [1, min(P_i)] and at most 64.  `tune` may be NULL. */
AKNN_API int aknn_forward(const aknn_dims* dims, const aknn_tune* tune, const void* q, const void* p,
                 const void* f, const void* sigma, const void* R, void* f_out, void* w_out,
                 void* dists, int64_t* idxs, void* workspace, size_t workspace_bytes,
                 aknn_cache* cache, void* stream);

/* Allocation-free forward (no counterpart in the reference: for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same as aknn_forward, with the same results, but without a
 * workspace argument: the library's own per-device scratch array of
 * AKNN_NOALLOC_SCRATCH_BYTES (in its code object, held until the process ends) takes the
 * workspace's place, and no call allocates, frees or runs an allocating routine.  Where
 * the tiled search's default split walk for the shape would not fit, a smaller split
 * (same results) is taken; a batch that needs more scratch than that even unsplit fails
 * with AKNN_E_UNSUPPORTED and the sizes in the message.  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (different streams, or threads)
 * race and must be serialised by the caller. */
#define AKNN_NOALLOC_SCRATCH_BYTES (64ull << 20)
AKNN_API int aknn_forward_noalloc(const aknn_dims* dims, const aknn_tune* tune, const void* q, const void* p,
                                  const void* f, const void* sigma, const void* R, void* f_out, void* w_out,
                                  void* dists, int64_t* idxs, aknn_cache* cache, void* stream);

/* Batched copies by a kernel (no counterpart in the reference; for a serving loop that
 * moves query rows in and results out without the runtime's copy engine path).  Each
 * item copies `bytes` from src to dst; either side may be device memory or mapped pinned
 * host memory (the device address from aknn_host_device_pointer).  One launch on
 * `stream` for up to AKNN_COPY_MAX_ITEMS items; items must not overlap. */
typedef struct aknn_copy_item {
  void* dst;
  const void* src;
  uint64_t bytes;
} aknn_copy_item;
#define AKNN_COPY_MAX_ITEMS 8
AKNN_API int aknn_copy(const aknn_copy_item* items, int n_items, void* stream);
/* Device address of pinned host memory allocated with hipHostMalloc (such as torch's
 * pin_memory tensors), for aknn_copy. */
AKNN_API int aknn_host_device_pointer(void* host, void** device);

/* Neighbour search only (idxs + dists).  Used by the aggregation-free paths. */
AKNN_API int aknn_knn(const aknn_dims* dims, const aknn_tune* tune, const void* q, const void* p,
             const void* sigma, const void* R, void* dists, int64_t* idxs, void* workspace,
             size_t workspace_bytes, void* stream);

/* Aggregation given neighbour indices (no search): recomputes dists with the same key
 * expression as the search (same bits), then w_out and f_out.  Replaces
 * knn_precomputed_aggregate_aniso_forward (csrc/knn_aggregate.cpp:76-99,
 * csrc/forward.cu:398-450); also the second half of aknn_forward.  idxs entries
 * must lie in [0, P_i) of their cloud. */
AKNN_API int aknn_aggregate(const aknn_dims* dims, const void* q, const void* p, const void* f,
                            const void* sigma, const void* R, const int64_t* idxs, void* f_out,
                            void* w_out, void* dists, void* stream);

/* Full second order (no counterpart in the reference, which drops these terms at
 * cuda_knn_aggregate/__init__.py:271-289): the backward of aknn_backward's map
 * (grad_f_out, q, p, f, sigma, R) -> (dq, dp, df, dsigma, dR) for an upstream on every
 * output: grad_grad_q [B,Q,3], grad_grad_p [B,P,3], grad_grad_f [B,P,F],
 * grad_grad_sigma [B,P,3], grad_grad_R [B,P,3,3], each optional (NULL = zero; not all
 * NULL).  Outputs as aknn_backward2.  With only grad_grad_q it equals aknn_backward2 up
 * to rounding.  dq always comes from the point kernels, so the workspace is
 * aknn_backward_workspace_size(dims, needs | AKNN_SECOND_FULL) and the CSR is required
 * whenever dq or a point output is requested. */
AKNN_API int aknn_backward2_full(const aknn_dims* dims, uint32_t needs, const void* grad_grad_q,
                                 const void* grad_grad_p, const void* grad_grad_f,
                                 const void* grad_grad_sigma, const void* grad_grad_R,
                                 const void* grad_f_out, const void* q, const void* p, const void* f,
                                 const void* sigma, const void* R, const int64_t* idxs,
                                 const int32_t* csr_offsets, const int32_t* csr_entries,
                                 const aknn_cache* cache, void* workspace, size_t workspace_bytes,
                                 void* d_grad_f_out, void* dq, void* dp, void* df, void* dsigma,
                                 void* dR, void* stream);

/* The grad_w_out path (no counterpart in the reference, whose first-order kernel keeps
 * "+ grad_w_out" commented out at csrc/backward.cu:93-94, so a loss on w_out contributes
 * nothing there; the default path here does the same).  With grad_w_out [query rows, K]
 * (same dtype as the inputs; NULL = zero) the upstream of the weight w_k of slot k is
 * grad_f_out . f[n_k] + grad_w_out[k] instead of its first term: every dq, dp, dsigma and
 * dR term gains it, df does not (w_out does not depend on f).
 *   aknn_backward_w: as aknn_backward (broadcast upstream allowed).
 *   aknn_backward2_w: the second order of that map, as aknn_backward2 when grad_grad_p
 *   .. grad_grad_R are all NULL, else as aknn_backward2_full (materialised grad_f_out,
 *   workspace for needs | AKNN_SECOND_FULL); with AKNN_NEED_DGRAD_W_OUT in needs it also
 *   writes d_grad_w_out [query rows, K], the gradient with respect to grad_w_out. */
#define AKNN_NEED_DGRAD_W_OUT 256u
AKNN_API int aknn_backward_w(const aknn_dims* dims, uint32_t needs, const void* grad_f_out,
                             const void* grad_w_out, const void* q, const void* p, const void* f,
                             const void* sigma, const void* R, const int64_t* idxs,
                             const int32_t* csr_offsets, const int32_t* csr_entries, const aknn_cache* cache,
                             void* workspace, size_t workspace_bytes, void* dq, void* dp, void* df,
                             void* dsigma, void* dR, void* stream);
AKNN_API int aknn_backward2_w(const aknn_dims* dims, uint32_t needs, const void* grad_grad_q,
                              const void* grad_grad_p, const void* grad_grad_f, const void* grad_grad_sigma,
                              const void* grad_grad_R, const void* grad_f_out, const void* grad_w_out,
                              const void* q, const void* p, const void* f, const void* sigma, const void* R,
                              const int64_t* idxs, const int32_t* csr_offsets, const int32_t* csr_entries,
                              const aknn_cache* cache, void* workspace, size_t workspace_bytes,
                              void* d_grad_f_out, void* d_grad_w_out, void* dq, void* dp, void* df,
                              void* dsigma, void* dR, void* stream);

/* Table-lookup forward.  Replaces knn_lookup_aggregate_aniso_forward
 * (csrc/knn_aggregate.cpp:50-73; kernel csrc/forward.cu:113-183, launcher :341-395).  For
 * each query: n1 = its Euclidean nearest point (the lowest index reaching the minimum of
 * x^2+y^2+z^2, and 0 unless that minimum is below 1e10, as the reference's scan
 * started from 1e10 gives); neighbours idxs[q, k] = knn_table[n1, k] for k < K (table
 * order, not sorted); then dists = anisotropic D, w_out = exp(-D) and f_out as in
 * aknn_aggregate.  knn_table: int64 [P, table_k] shared by all B batches, entries in
 * [0, P) (not checked here); K <= table_k.  Dense batches only.  The backward passes
 * take its idxs like any other. */
AKNN_API size_t aknn_lookup_workspace_size(const aknn_dims* dims);
AKNN_API int aknn_lookup_forward(const aknn_dims* dims, const void* q, const void* p, const void* f,
                                 const void* sigma, const void* R, const int64_t* knn_table,
                                 int64_t table_k, void* f_out, void* w_out, void* dists, int64_t* idxs,
                                 void* workspace, size_t workspace_bytes, void* stream);

/* Isotropic forward.  Replaces knn_aggregate_forward (csrc/knn_aggregate.cpp:5-23;
 * kernel csrc/forward.cu:5-53, launcher :233-283): the K nearest points by Euclidean
 * distance ((key, index) order), dists = x^2+y^2+z^2, w_k = exp(-(dists_k * scale[n_k])),
 * f_out = sum_k w_k f[n_k], w_out[B, Q] (one value per query, the reference's [B, Q, 1]) =
 * sum_k w_k.  scale: one value per point row (the reference reads sigma[b, n, 0]).
 * dims->key is ignored.  No backward (the reference has none). */
AKNN_API size_t aknn_iso_workspace_size(const aknn_dims* dims);
AKNN_API int aknn_iso_forward(const aknn_dims* dims, const aknn_tune* tune, const void* q, const void* p,
                              const void* f, const void* scale, void* f_out, void* w_out, void* dists,
                              int64_t* idxs, void* workspace, size_t workspace_bytes, void* stream);

/* Point-side transpose of idxs (CSR over the B*P point rows, each row's entries
 * q*K+k in ascending order), needed by the point gradients of both backward
 * passes.  csr_offsets: int32 [B*P + 1]; csr_entries: int32 [B*Q*K]. */
AKNN_API size_t aknn_csr_workspace_size(const aknn_dims* dims);
AKNN_API int aknn_build_csr(const aknn_dims* dims, const int64_t* idxs, int32_t* csr_offsets,
                   int32_t* csr_entries, void* workspace, size_t workspace_bytes, void* stream);
/* The same transpose in a spatial numbering of the queries (no counterpart in the
 * reference): query_order [query rows] lists the query rows in the forward's spatial
 * order (aknn_cache.query_order); entry (q, k) is numbered (position of q in that
 * order)*K + k, each row's entries still in ascending q*K+k order (so the backward
 * sums, and their bits, are those of the plain transpose).  Given AKNN_CSR_QUERY_ORDER
 * in needs and the same cache, the backward passes read the query-side inputs from
 * copies in that order, where the queries of one point row lie close together.  Same
 * workspace as aknn_build_csr. */
#define AKNN_CSR_QUERY_ORDER 512u
AKNN_API int aknn_build_csr_ordered(const aknn_dims* dims, const int64_t* idxs, const int32_t* query_order,
                                    int32_t* csr_offsets, int32_t* csr_entries, void* workspace,
                                    size_t workspace_bytes, void* stream);

/* Device workspace of aknn_backward / aknn_backward2 for a `needs` mask (0 when none is
 * needed): with dq and point gradients both requested, the point kernels hand each
 * (query, slot) entry's dq share to a per-query sum instead of a second gather.  With
 * AKNN_GRAD_BROADCAST and dq requested: the dot records (dtype[point rows][16]),
 * unless the cache brings them.  With AKNN_CSR_QUERY_ORDER: also the ordered copies of
 * the query-side inputs the point kernels read. */
AKNN_API size_t aknn_backward_workspace_size(const aknn_dims* dims, uint32_t needs);

/* Dot records for a broadcast upstream (no counterpart in the reference): for every
 * point row n, 16 values {p0 p1 p2, R00..R22, s0 s1 s2, a} with
 * a = grad_rows[cloud(n)] . f[n] in the association the backward kernels use (same
 * bits); grad_rows [B, F], out dtype[point rows][16].  dq of a broadcast upstream reads
 * a neighbour's geometry and dot from one record.  For aknn_cache.point_dots. */
AKNN_API int aknn_point_dots(const aknn_dims* dims, const void* grad_rows, const void* f, const void* p,
                             const void* sigma, const void* R, void* out, void* stream);

/* First-order backward.  Replaces knn_aggregate_aniso_backward
 * (csrc/knn_aggregate.cpp:101-132).  Gaussian weights are recomputed from idxs;
 * grad_w_out is not an input (the reference ignores it, csrc/backward.cu:93-94).
 * Outputs selected by `needs` are fully overwritten; others may be NULL.  The
 * point outputs need the CSR from aknn_build_csr.  grad_f_out is [B, Q, F], or [B, F]
 * with AKNN_GRAD_BROADCAST in needs. */
AKNN_API int aknn_backward(const aknn_dims* dims, uint32_t needs, const void* grad_f_out, const void* q,
                  const void* p, const void* f, const void* sigma, const void* R,
                  const int64_t* idxs, const int32_t* csr_offsets, const int32_t* csr_entries,
                  const aknn_cache* cache, void* workspace, size_t workspace_bytes, void* dq,
                  void* dp, void* df, void* dsigma, void* dR, void* stream);

/* Second-order backward of <grad_grad_q, dq>.  Replaces knn_aggregate_aniso_backward_2nd
 * (csrc/knn_aggregate.cpp:135-172); like the reference only grad_grad_q is an input
 * (cuda_knn_aggregate/__init__.py:260-289).  grad_f_out as in aknn_backward (the
 * broadcast bit applies here too; d_grad_f_out stays [B, Q, F]). */
AKNN_API int aknn_backward2(const aknn_dims* dims, uint32_t needs, const void* grad_grad_q,
                   const void* grad_f_out, const void* q, const void* p, const void* f,
                   const void* sigma, const void* R, const int64_t* idxs,
                   const int32_t* csr_offsets, const int32_t* csr_entries,
                   const aknn_cache* cache, void* workspace, size_t workspace_bytes,
                   void* d_grad_f_out, void* dq, void* dp, void* df, void* dsigma, void* dR,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AKNN_H_ */
