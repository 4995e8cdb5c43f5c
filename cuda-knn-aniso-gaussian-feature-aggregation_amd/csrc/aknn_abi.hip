// C ABI entry points (declared in include/aknn.h).
//
// Each entry validates the problem description, then enqueues the kernels on the
// caller's stream.  Nothing here allocates device memory or synchronises.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "aknn_common.h"
#include "aknn_internal.h"

namespace {

thread_local std::string g_last_error;

// ------------------------------------------------------------------ tracing
// When enabled, every library launch is bracketed by two hipEvents recorded on the
// launch stream; aknn_trace_collect waits for them and reports per-launch times.
// Off by default (zero cost); not for use under stream capture.
struct TraceRec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex g_trace_mu;
bool g_trace_on = false;
std::vector<TraceRec> g_trace;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

class Span {
 public:
  Span(const char* name, hipStream_t st) : st_(st) {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    if (!g_trace_on) return;
    rec_.name = name;
    rec_.a = pool_event();
    rec_.b = pool_event();
    if (rec_.a && rec_.b) {
      on_ = hipEventRecord(rec_.a, st_) == hipSuccess;  // tracing only: a failed record drops the span
    }
    if (!on_) {  // the events go back to the pool
      if (rec_.a) g_event_pool.push_back(rec_.a);
      if (rec_.b) g_event_pool.push_back(rec_.b);
    }
  }
  ~Span() {
    if (!on_) return;
    if (hipEventRecord(rec_.b, st_) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    g_trace.push_back(rec_);
  }

 private:
  hipStream_t st_;
  TraceRec rec_{};
  bool on_ = false;
};

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  if (e == hipErrorNotSupported)
    return fail(AKNN_E_UNSUPPORTED,
                "%s: feature width not supported (F <= 128, or F a multiple of 4 up to 512 "
                "in float32 / of 2 up to 256 in float64)",
                where);
  return fail(AKNN_E_LAUNCH, "%s: %s", where, hipGetErrorString(e));
}

int64_t elem_size(const aknn_dims* d) { return d->dtype == AKNN_F64 ? 8 : 4; }

bool ragged(const aknn_dims* d) { return d->q_off != nullptr; }

// Smallest per-cloud point count (K must not exceed it).
int64_t min_points(const aknn_dims* d) {
  if (!ragged(d)) return d->P;
  int64_t m = INT64_MAX;
  for (int64_t b = 0; b < d->B; ++b) m = std::min(m, d->p_off_host[b + 1] - d->p_off_host[b]);
  return m;
}

int64_t query_rows(const aknn_dims* d) { return ragged(d) ? d->Q : d->B * d->Q; }
int64_t point_rows(const aknn_dims* d) { return ragged(d) ? d->P : d->B * d->P; }

int check_dims(const aknn_dims* d, bool need_k) {
  if (!d) return fail(AKNN_E_INVALID, "dims is NULL");
  if (d->dtype != AKNN_F32 && d->dtype != AKNN_F64)
    return fail(AKNN_E_INVALID, "dtype must be AKNN_F32 or AKNN_F64 (got %d)", d->dtype);
  if (d->key != AKNN_KEY_ANISO && d->key != AKNN_KEY_EUCLIDEAN)
    return fail(AKNN_E_INVALID, "unknown key mode %d", d->key);
  if (d->B < 0 || d->Q < 0 || d->P < 0 || d->F < 1)
    return fail(AKNN_E_INVALID, "bad sizes B=%lld Q=%lld P=%lld F=%lld", (long long)d->B,
                (long long)d->Q, (long long)d->P, (long long)d->F);
  // The gather kernels hold a feature row in registers: 16 lanes x up to 8 chunks of the
  // widest vector (at most 16 B) that divides F.
  const int64_t fvec = (d->dtype == AKNN_F32 && d->F % 4 == 0) ? 4 : (d->F % 2 == 0 ? 2 : 1);
  if (d->F > 128 * fvec)
    return fail(AKNN_E_UNSUPPORTED, "F must be <= %lld for this F (128 x the widest vector of <= 16 B "
                "dividing F; got %lld)", (long long)(128 * fvec), (long long)d->F);
  const bool rq = d->q_off != nullptr, rp = d->p_off != nullptr;
  if (rq != rp || (rq && (!d->q_off_host || !d->p_off_host)))
    return fail(AKNN_E_INVALID, "ragged batches need device and host offsets for q and p");
  if (rq) {
    if (d->q_off_host[0] != 0 || d->p_off_host[0] != 0 || d->q_off_host[d->B] != d->Q ||
        d->p_off_host[d->B] != d->P)
      return fail(AKNN_E_INVALID, "offsets must start at 0 and end at the total row counts");
    for (int64_t b = 0; b < d->B; ++b)
      if (d->q_off_host[b + 1] < d->q_off_host[b] || d->p_off_host[b + 1] < d->p_off_host[b])
        return fail(AKNN_E_INVALID, "offsets must be non-decreasing");
  }
  if (need_k) {
    if (d->K < 1) return fail(AKNN_E_INVALID, "K must be >= 1 (got %lld)", (long long)d->K);
    if (d->K > 64) return fail(AKNN_E_UNSUPPORTED, "K must be <= 64 (got %lld)", (long long)d->K);
    if (d->B > 0 && d->K > min_points(d))
      return fail(AKNN_E_INVALID, "K (%lld) must not exceed the number of points (%lld)",
                  (long long)d->K, (long long)min_points(d));
  }
  if (point_rows(d) >= (int64_t(1) << 31) || query_rows(d) * std::max<int64_t>(d->K, 1) >= (int64_t(1) << 31))
    return fail(AKNN_E_UNSUPPORTED, "batch too large (point rows and query rows * K must be < 2^31)");
  return AKNN_OK;
}

aknn::Clouds clouds_of(const aknn_dims* d) {
  aknn::Clouds c;
  c.q_off = d->q_off;
  c.p_off = d->p_off;
  c.B = d->B;
  c.Q = d->Q;
  c.P = d->P;
  return c;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Launch shape of the top-K scan: queries per lane and waves splitting the points.
void pick_knn_shape(const aknn_dims* d, const aknn_tune* t, int* qpt, int* split) {
  int qq = t && t->queries_per_lane > 0 ? t->queries_per_lane : 1;
  int ss = t && t->point_split > 0 ? t->point_split : 0;
  if (qq != 1 && qq != 2 && qq != 4) qq = 1;
  if (ss == 0) {
    // enough waves to cover the 1024 SIMDs several times over
    const int64_t waves = (query_rows(d) + 64 * qq - 1) / (64 * qq);
    const int64_t pmin = ragged(d) ? std::max<int64_t>(1, d->P / std::max<int64_t>(1, d->B)) : d->P;
    ss = 1;
    while (ss < 8 && waves * ss < 4096 && pmin / (2 * ss) >= 4 * d->K) ss *= 2;
  }
  if (ss != 1 && ss != 2 && ss != 4 && ss != 8) ss = 1;
  *qpt = qq;
  *split = ss;
}

int64_t knn_tiles(const aknn_dims* d, int qpt) {
  int64_t n = 0;
  for (int64_t b = 0; b < d->B; ++b) {
    const int64_t qb = d->q_off_host[b + 1] - d->q_off_host[b];
    n += (qb + 64 * qpt - 1) / (64 * qpt);
  }
  return n;
}

// Ragged tile table: tile_prefix[b] = first tile of cloud b (B+1 entries), one thread.
__global__ void tile_prefix_kernel(const int64_t* __restrict__ q_off, int64_t B, int64_t per,
                                   int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  for (int64_t b = 0; b < B; ++b) {
    out[b] = acc;
    acc += (q_off[b + 1] - q_off[b] + per - 1) / per;
  }
  out[B] = acc;
}

size_t rec_bytes(const aknn_dims* d) { return align256(static_cast<size_t>(point_rows(d)) * 16 * elem_size(d)); }

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr int64_t kSplitBlocks = 16384;

// Super-tiles of the batch at `per` tiles each.
int64_t super_count(const aknn_dims* d, int64_t per) {
  if (!ragged(d)) return d->B * ceil_div(ceil_div(d->P, aknn::kTile), per);
  int64_t n = 0;
  for (int64_t b = 0; b < d->B; ++b) n += ceil_div(ceil_div(d->p_off_host[b + 1] - d->p_off_host[b], aknn::kTile), per);
  return n;
}

void set_super(aknn::PrunedLaunch& L, const aknn_dims* d, int per) {
  L.super_tiles = per;
  L.nsuper = super_count(d, per);
}

// Tile geometry and occupancy from a tune (results never depend on them).  A prebuilt
// point index fixes the defaults it was built with; sub-tile sizes and occupancy
// targets other than the defaults exist for the FP32 search with 9 to 16 neighbours.
void apply_geometry(aknn::PrunedLaunch& L, const aknn_dims* d, const aknn_tune* t) {
  if (!t || t->index) return;
  if (t->super_tiles == 8 || t->super_tiles == 16 || t->super_tiles == 32) set_super(L, d, t->super_tiles);
  const bool variants = d->dtype == AKNN_F32 && d->K > 8 && d->K <= 16;
  if (variants && (t->sub_records == 4 || t->sub_records == 8 || t->sub_records == 16)) L.sub_rec = t->sub_records;
  if (variants && L.sub_rec == aknn::kSubRecDefault && (t->waves_per_eu == 4 || t->waves_per_eu == 8))
    L.waves = t->waves_per_eu;
  if (t->mu_classes == 1 || t->mu_classes == 2 || t->mu_classes == 4 || t->mu_classes == 8)
    L.mu_classes = t->mu_classes;
}

aknn::PrunedLaunch pruned_launch(const aknn_dims* d) {
  aknn::PrunedLaunch L{};
  L.dtype = d->dtype;
  L.key = d->key;
  L.cl = clouds_of(d);
  L.qrows = query_rows(d);
  L.prows = point_rows(d);
  L.K = d->K;
  if (ragged(d)) {
    for (int64_t b = 0; b < d->B; ++b) {
      const int64_t pb = d->p_off_host[b + 1] - d->p_off_host[b];
      const int64_t tiles = ceil_div(pb, aknn::kTile);
      L.ntiles += tiles;
      L.nqtiles += ceil_div(d->q_off_host[b + 1] - d->q_off_host[b], aknn::kQueryTile);
      L.max_points = std::max(L.max_points, pb);
    }
  } else {
    L.ntiles = d->B * ceil_div(d->P, aknn::kTile);
    L.nqtiles = d->B * ceil_div(d->Q, aknn::kQueryTile);
    L.max_points = d->P;
  }
  L.sub_rec = aknn::kSubRecDefault;
  L.waves = 0;
  L.mu_classes = aknn::kMuClassesDefault;
  // Ragged batches: resident blocks over the (cloud, query tile) work list, measured
  // faster than one block per query tile there (uneven clouds); dense batches keep the
  // one-block-per-tile grid, measured faster on clustered clouds.
  L.persistent = ragged(d) ? 1 : 0;
  L.nsuper_alloc = super_count(d, aknn::kSuperMin);
  set_super(L, d, aknn::kSuper);
  // Few query tiles (small or scattered query sets) cannot fill the device and the
  // slowest tile's walk sets the kernel time: split every tile's walk over up to 64
  // blocks with at least 8 point tiles per block and cloud, and at most kSplitBlocks
  // blocks in all (the workspace is sized for that largest split, ways_max).  The
  // default follows ways^2 * query tiles ~ 65536, the measured optimum on 64K-point
  // clouds from 4 to 4096 query tiles; a tuning value may pick anything up to ways_max.
  const int64_t tiles_per_cloud = L.ntiles / std::max<int64_t>(1, d->B);
  int w = 1;
  while (w < 64 && L.nqtiles * (2 * w) <= kSplitBlocks && tiles_per_cloud >= 8 * (2 * w)) w *= 2;
  L.ways_max = w;
  L.ways = 1;
  while (2 * L.ways <= L.ways_max && L.nqtiles * (4 * L.ways * L.ways) <= 65536) L.ways *= 2;
  return L;
}

// Default search: the tiled search once clouds hold enough points for pruning to pay
// (measured crossover ~8K points per cloud), the brute-force scan below that.
int pick_search(const aknn_dims* d, const aknn_tune* t) {
  if (t && t->index) return AKNN_SEARCH_TILES;  // a prebuilt point index is for the tiled search
  if (t && t->search == AKNN_SEARCH_WAVE) return AKNN_SEARCH_TILES;  // same index, per-query walk
  if (t && (t->search == AKNN_SEARCH_BRUTE || t->search == AKNN_SEARCH_TILES)) return t->search;
  const int64_t mean_p = d->B > 0 ? point_rows(d) / d->B : 0;
  return mean_p >= 8192 ? AKNN_SEARCH_TILES : AKNN_SEARCH_BRUTE;
}

// Within the tiled search: the per-query walk when the queries are sparse in their
// clouds -- at most one per point of the clouds that have queries -- where the 64
// queries of a query tile share few point tiles.  Measured on 64K-point clustered
// clouds: 5x faster at 256 queries, 1.3x at 64K, 2x slower at 256K.

bool pick_wave(const aknn_dims* d, const aknn_tune* t) {
  if (t && t->search == AKNN_SEARCH_WAVE) return true;
  if (t && t->search == AKNN_SEARCH_TILES) return false;
  int64_t pts = 0;
  if (ragged(d)) {
    for (int64_t b = 0; b < d->B; ++b)
      if (d->q_off_host[b + 1] > d->q_off_host[b]) pts += d->p_off_host[b + 1] - d->p_off_host[b];
  } else if (d->Q > 0) {
    pts = d->B * d->P;
  }
  return query_rows(d) <= pts;
}

// Neighbour search.  On return *qorder / *porder point at spatial orders of the query
// and point rows inside the workspace (tiled search) or are NULL (brute scan).
int run_knn(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p,
            const void* sigma, const void* R, void* dists, int64_t* idxs, void* ws,
            size_t ws_bytes, hipStream_t st, const uint32_t** qorder = nullptr,
            const uint32_t** porder = nullptr) {
  if (qorder) *qorder = nullptr;
  if (porder) *porder = nullptr;
  if (pick_search(d, tune) == AKNN_SEARCH_TILES) {
    aknn::PrunedLaunch PL = pruned_launch(d);
    if (PL.ntiles >= (int64_t(1) << 30))
      return fail(AKNN_E_UNSUPPORTED, "%lld point tiles exceed the tiled search's 2^30 (use the brute scan)",
                  static_cast<long long>(PL.ntiles));
    PL.q = q;
    PL.p = p;
    PL.sigma = sigma;
    PL.R = R;
    PL.dists = dists;
    PL.idxs = idxs;
    PL.ws = ws;
    PL.ws_bytes = ws_bytes;
    PL.stats = tune ? static_cast<unsigned long long*>(tune->stats) : nullptr;
    PL.index = tune ? tune->index : nullptr;
    PL.wave = pick_wave(d, tune) ? 1 : 0;
    apply_geometry(PL, d, tune);
    if (tune && tune->persistent != 0) PL.persistent = tune->persistent > 0;
    if (tune && tune->point_split > 0) {
      int w = 1;
      while (2 * w <= tune->point_split && 2 * w <= PL.ways_max) w *= 2;
      PL.ways = w;
    }
    hipError_t e;
    {
      Span sp("knn_tiles_prep", st);
      e = aknn::launch_knn_pruned_prep(PL, st);
    }
    if (e != hipSuccess) return hip_fail(e, "knn_tiles_prep");
    {
      Span sp(PL.wave ? "knn_wave" : "knn_tiles", st);
      e = aknn::launch_knn_pruned_search(PL, st);
    }
    if (e != hipSuccess) return hip_fail(e, "knn_tiles");
    if (qorder) *qorder = aknn::pruned_query_order(PL);
    if (porder) *porder = aknn::pruned_point_order(PL);
    return AKNN_OK;
  }
  char* base = static_cast<char*>(ws);
  void* rec = base;
  hipError_t e;
  {
    Span sp("pack_points", st);
    e = d->dtype == AKNN_F32 ? aknn::launch_pack_points<float>(p, sigma, R, rec, point_rows(d), st)
                             : aknn::launch_pack_points<double>(p, sigma, R, rec, point_rows(d), st);
  }
  if (e != hipSuccess) return hip_fail(e, "pack_points");
  aknn::KnnLaunch L{};
  L.dtype = d->dtype;
  L.key = d->key;
  L.cl = clouds_of(d);
  L.K = d->K;
  pick_knn_shape(d, tune, &L.qpt, &L.split);
  L.q = q;
  L.rec = rec;
  L.dists = dists;
  L.idxs = idxs;
  L.tiles = nullptr;
  L.ntiles = 0;
  if (ragged(d)) {
    int64_t* prefix = reinterpret_cast<int64_t*>(base + rec_bytes(d));
    hipLaunchKernelGGL(tile_prefix_kernel, dim3(1), dim3(64), 0, st, d->q_off, d->B,
                       int64_t(64) * L.qpt, prefix);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tile_prefix");
    L.tiles = prefix;
    L.ntiles = knn_tiles(d, L.qpt);
    if (L.ntiles == 0) return AKNN_OK;
  }
  {
    Span sp("knn_topk", st);
    e = aknn::launch_knn(L, st);
  }
  if (e != hipSuccess) return hip_fail(e, "knn_topk");
  return AKNN_OK;
}

}  // namespace

extern "C" {

int aknn_abi_version(void) { return AKNN_ABI_VERSION; }

void aknn_trace_enable(int on) {
  std::lock_guard<std::mutex> lk(g_trace_mu);
  g_trace_on = on != 0;
}

int aknn_trace_collect(char* names, size_t names_bytes, float* ms, int max_records) {
  std::vector<TraceRec> recs;
  {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    recs.swap(g_trace);
  }
  int n = 0;
  size_t used = 0;
  for (const TraceRec& r : recs) {
    float t = -1.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) t = -1.f;
    if (n < max_records && ms && names) {
      const size_t len = strlen(r.name) + 1;
      if (used + len <= names_bytes) {
        memcpy(names + used, r.name, len);
        used += len;
        ms[n++] = t;
      }
    }
  }
  std::lock_guard<std::mutex> lk(g_trace_mu);
  for (const TraceRec& r : recs) {
    g_event_pool.push_back(r.a);
    g_event_pool.push_back(r.b);
  }
  return n;
}

const char* aknn_last_error(void) { return g_last_error.c_str(); }

size_t aknn_forward_workspace_size(const aknn_dims* d) {
  if (check_dims(d, false) != AKNN_OK) return 0;
  const size_t brute = rec_bytes(d) + align256(static_cast<size_t>(d->B + 1) * 8);
  size_t tiles = 0;
  const hipError_t e = aknn::pruned_workspace(pruned_launch(d), &tiles);
  if (e != hipSuccess) {
    hip_fail(e, "forward_workspace_size");
    return 0;
  }
  return std::max(brute, tiles);
}

int aknn_knn(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p,
             const void* sigma, const void* R, void* dists, int64_t* idxs, void* ws,
             size_t ws_bytes, void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (query_rows(d) == 0 || d->B == 0) return AKNN_OK;
  if (!q || !p || !sigma || !R || !dists || !idxs) return fail(AKNN_E_INVALID, "null tensor pointer");
  if (!ws || ws_bytes < aknn_forward_workspace_size(d))
    return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes,
                aknn_forward_workspace_size(d));
  return run_knn(d, tune, q, p, sigma, R, dists, idxs, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// Search + aggregation into a workspace already checked to suffice for this tune.
static int forward_run(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p, const void* f,
                       const void* sigma, const void* R, void* f_out, void* w_out, void* dists, int64_t* idxs,
                       void* ws, size_t ws_bytes, aknn_cache* cache, void* stream) {
  int rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t *qord = nullptr, *pord = nullptr;
  rc = run_knn(d, tune, q, p, sigma, R, dists, idxs, ws, ws_bytes, st, &qord, &pord);
  if (rc) return rc;
  // cache for the backward passes (and the aggregation below)
  const void* rec = nullptr;
  const int32_t* order = reinterpret_cast<const int32_t*>(qord);
  if (cache) {
    const hipError_t e =
        d->dtype == AKNN_F32
            ? aknn::launch_cache_fill<float>(qord, cache->query_order, query_rows(d), pord, cache->point_order,
                                             point_rows(d), p, sigma, R, cache->point_records, st)
            : aknn::launch_cache_fill<double>(qord, cache->query_order, query_rows(d), pord, cache->point_order,
                                              point_rows(d), p, sigma, R, cache->point_records, st);
    if (e != hipSuccess) return hip_fail(e, "cache");
    rec = cache->point_records;
  }
  aknn::QueryLaunch L{};
  L.dtype = d->dtype;
  L.key = d->key;
  L.cl = clouds_of(d);
  L.rows = query_rows(d);
  L.F = d->F;
  L.K = d->K;
  L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
  L.idxs = idxs;
  L.dists = dists;
  L.dists_out = dists;  // Euclidean ranking: overwritten with the anisotropic D
  L.f_out = f_out;
  L.w_out = w_out;
  L.order = order;
  L.rec = rec;
  hipError_t e;
  {
    Span sp("aggregate", st);
    e = aknn::launch_aggregate(L, st);
  }
  if (e != hipSuccess) return hip_fail(e, "aggregate");
  return AKNN_OK;
}

extern "C" {

int aknn_forward(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p,
                 const void* f, const void* sigma, const void* R, void* f_out, void* w_out,
                 void* dists, int64_t* idxs, void* ws, size_t ws_bytes, aknn_cache* cache,
                 void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (query_rows(d) == 0 || d->B == 0) return AKNN_OK;
  if (!q || !p || !sigma || !R || !f || !f_out || !w_out || !dists || !idxs)
    return fail(AKNN_E_INVALID, "null tensor pointer");
  if (!ws || ws_bytes < aknn_forward_workspace_size(d))
    return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes,
                aknn_forward_workspace_size(d));
  return forward_run(d, tune, q, p, f, sigma, R, f_out, w_out, dists, idxs, ws, ws_bytes, cache, stream);
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free forward
// Device memory the allocation-free forward uses in place of a caller's workspace: one
// copy per GPU, in the library's code object, held until the process ends.  Host code
// never passes `scratch` itself (in host code it is only a stand-in for the symbol):
// each device's address comes from hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = size_t(AKNN_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch` (looked up on the device's first call).
int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "forward_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(AKNN_E_UNSUPPORTED, "forward_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                 "forward_noalloc: hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return AKNN_OK;
}

// Workspace bytes the forward needs with the tiled search split `ways` ways (or the brute
// scan's, for the shapes that take it).
size_t forward_need(const aknn_dims* d, const aknn_tune* tune, int ways) {
  const size_t brute = rec_bytes(d) + align256(static_cast<size_t>(d->B + 1) * 8);
  if (pick_search(d, tune) != AKNN_SEARCH_TILES) return brute;
  aknn::PrunedLaunch PL = pruned_launch(d);
  PL.ways = ways;
  PL.ways_max = ways;
  PL.index = tune ? tune->index : nullptr;  // a prebuilt point index is not carved from the workspace
  size_t bytes = 0;
  if (aknn::pruned_workspace(PL, &bytes) != hipSuccess) return SIZE_MAX;
  return bytes;
}
}  // namespace

extern "C" {

int aknn_forward_noalloc(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p, const void* f,
                         const void* sigma, const void* R, void* f_out, void* w_out, void* dists, int64_t* idxs,
                         aknn_cache* cache, void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (query_rows(d) == 0 || d->B == 0) return AKNN_OK;
  if (!q || !p || !sigma || !R || !f || !f_out || !w_out || !dists || !idxs)
    return fail(AKNN_E_INVALID, "null tensor pointer");
  // The tiled search's split walk keeps per-way lists; where the library's split for
  // the shape does not fit the scratch, take the largest split that does (every split
  // gives the same results).
  aknn_tune t = tune ? *tune : aknn_tune{};
  int ways = 1;
  if (pick_search(d, tune) == AKNN_SEARCH_TILES) {
    const aknn::PrunedLaunch PL = pruned_launch(d);
    ways = PL.ways;
    if (tune && tune->point_split > 0) {
      int w = 1;
      while (2 * w <= tune->point_split && 2 * w <= PL.ways_max) w *= 2;
      ways = w;
    }
    while (ways > 1 && forward_need(d, tune, ways) > kScratchBytes) ways /= 2;
    t.point_split = ways;
  }
  const size_t need = forward_need(d, tune, ways);
  if (need > kScratchBytes)
    return fail(AKNN_E_UNSUPPORTED,
                "forward_noalloc: this batch needs %zu bytes of scratch, more than the %zu the allocation-free "
                "entry holds (use aknn_forward with a workspace)", need, kScratchBytes);
  void* ws = nullptr;
  rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  return forward_run(d, &t, q, p, f, sigma, R, f_out, w_out, dists, idxs, ws, kScratchBytes, cache, stream);
}

}  // extern "C"

// ------------------------------------------------------------------ batched copies
namespace {
struct CopyArgs {
  aknn_copy_item it[AKNN_COPY_MAX_ITEMS];
};

// blockIdx.y = item; the blocks of an item stride over it in 16-byte chunks when both
// ends are 16-byte aligned, then over the remaining bytes one at a time.  Every access
// stays inside [0, bytes) of its item.
__global__ __launch_bounds__(256) void copy_items_kernel(CopyArgs a) {
  const aknn_copy_item it = a.it[blockIdx.y];
  char* __restrict__ d = static_cast<char*>(it.dst);
  const char* __restrict__ s = static_cast<const char*>(it.src);
  const uint64_t n = it.bytes;
  const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t done = 0;
  if ((reinterpret_cast<uintptr_t>(d) | reinterpret_cast<uintptr_t>(s)) % 16 == 0) {
    const uint64_t n16 = n / 16;
    for (uint64_t k = tid; k < n16; k += stride)
      reinterpret_cast<uint4*>(d)[k] = reinterpret_cast<const uint4*>(s)[k];
    done = n16 * 16;
  }
  for (uint64_t k = done + tid; k < n; k += stride) d[k] = s[k];
}
}  // namespace

extern "C" {

int aknn_copy(const aknn_copy_item* items, int n_items, void* stream) {
  if (n_items < 0 || n_items > AKNN_COPY_MAX_ITEMS)
    return fail(AKNN_E_INVALID, "aknn_copy: 0 to %d items per call (got %d)", AKNN_COPY_MAX_ITEMS, n_items);
  if (n_items == 0) return AKNN_OK;
  if (!items) return fail(AKNN_E_INVALID, "aknn_copy: items is NULL");
  CopyArgs a{};
  uint64_t most = 0;
  for (int i = 0; i < n_items; ++i) {
    if (items[i].bytes && (!items[i].dst || !items[i].src)) return fail(AKNN_E_INVALID, "aknn_copy: null pointer");
    a.it[i] = items[i];
    most = std::max<uint64_t>(most, items[i].bytes);
  }
  if (most == 0) return AKNN_OK;
  const uint64_t chunks = (most + 15) / 16;
  const unsigned bx = static_cast<unsigned>(std::min<uint64_t>(1024, std::max<uint64_t>(1, (chunks + 255) / 256)));
  hipStream_t st = static_cast<hipStream_t>(stream);
  {
    Span sp("copy", st);
    hipLaunchKernelGGL(copy_items_kernel, dim3(bx, n_items), dim3(256), 0, st, a);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "copy");
  return AKNN_OK;
}

int aknn_host_device_pointer(void* host, void** device) {
  if (!host || !device) return fail(AKNN_E_INVALID, "aknn_host_device_pointer: null pointer");
  *device = nullptr;
  const hipError_t e = hipHostGetDevicePointer(device, host, 0);
  if (e != hipSuccess) return hip_fail(e, "hipHostGetDevicePointer (memory not from hipHostMalloc?)");
  return AKNN_OK;
}

int aknn_aggregate(const aknn_dims* d, const void* q, const void* p, const void* f,
                   const void* sigma, const void* R, const int64_t* idxs, void* f_out,
                   void* w_out, void* dists, void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (query_rows(d) == 0 || d->B == 0) return AKNN_OK;
  if (!q || !p || !f || !sigma || !R || !idxs || !f_out || !w_out || !dists)
    return fail(AKNN_E_INVALID, "null tensor pointer");
  aknn::QueryLaunch L{};
  L.dtype = d->dtype;
  L.key = d->key;
  L.cl = clouds_of(d);
  L.rows = query_rows(d);
  L.F = d->F;
  L.K = d->K;
  L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
  L.idxs = idxs;
  L.dists = nullptr;
  L.dists_out = dists;
  L.f_out = f_out;
  L.w_out = w_out;
  hipError_t e;
  {
    Span sp("aggregate_precomputed", static_cast<hipStream_t>(stream));
    e = aknn::launch_aggregate(L, static_cast<hipStream_t>(stream));
  }
  if (e != hipSuccess) return hip_fail(e, "aggregate");
  return AKNN_OK;
}

size_t aknn_index_size(const aknn_dims* d) {
  if (check_dims(d, false) != AKNN_OK) return 0;
  return aknn::pruned_index_bytes(pruned_launch(d));
}

size_t aknn_index_workspace_size(const aknn_dims* d) {
  if (check_dims(d, false) != AKNN_OK) return 0;
  return aknn::pruned_index_build_ws(pruned_launch(d));
}

int aknn_index_build(const aknn_dims* d, const void* p, const void* sigma, const void* R, void* index,
                     size_t index_bytes, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_dims(d, false);
  if (rc) return rc;
  if (point_rows(d) == 0 || d->B == 0) return fail(AKNN_E_INVALID, "no points to index");
  if (!p || !index || (d->key == AKNN_KEY_ANISO && (!sigma || !R)))
    return fail(AKNN_E_INVALID, "null pointer");
  aknn::PrunedLaunch PL = pruned_launch(d);
  if (PL.ntiles >= (int64_t(1) << 30)) return fail(AKNN_E_UNSUPPORTED, "too many point tiles");
  const size_t need = aknn::pruned_index_bytes(PL);
  if (index_bytes < need) return fail(AKNN_E_WORKSPACE, "index buffer too small (%zu < %zu bytes)", index_bytes, need);
  const size_t wneed = aknn::pruned_index_build_ws(PL);
  if (!ws || ws_bytes < wneed) return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes, wneed);
  PL.p = p;
  PL.sigma = sigma;
  PL.R = R;
  PL.index = index;
  PL.ws = ws;
  PL.ws_bytes = ws_bytes;
  hipError_t e;
  {
    Span sp("index_build", static_cast<hipStream_t>(stream));
    e = aknn::launch_pruned_index_build(PL, static_cast<hipStream_t>(stream));
  }
  if (e != hipSuccess) return hip_fail(e, "index_build");
  return AKNN_OK;
}

size_t aknn_iso_workspace_size(const aknn_dims* d) {
  if (!d) return 0;
  aknn_dims e = *d;
  e.key = AKNN_KEY_EUCLIDEAN;
  return aknn_forward_workspace_size(&e);
}

int aknn_iso_forward(const aknn_dims* d, const aknn_tune* tune, const void* q, const void* p,
                     const void* f, const void* scale, void* f_out, void* w_out, void* dists,
                     int64_t* idxs, void* ws, size_t ws_bytes, void* stream) {
  if (!d) return fail(AKNN_E_INVALID, "dims is NULL");
  aknn_dims e = *d;
  e.key = AKNN_KEY_EUCLIDEAN;
  int rc = check_dims(&e, true);
  if (rc) return rc;
  if (query_rows(&e) == 0 || e.B == 0) return AKNN_OK;
  if (!q || !p || !f || !scale || !f_out || !w_out || !dists || !idxs)
    return fail(AKNN_E_INVALID, "null tensor pointer");
  const size_t need = aknn_iso_workspace_size(&e);
  if (need == 0) return AKNN_E_LAUNCH;
  if (!ws || ws_bytes < need) return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes, need);
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = run_knn(&e, tune, q, p, nullptr, nullptr, dists, idxs, ws, ws_bytes, st);
  if (rc) return rc;
  aknn::QueryLaunch L{};
  L.dtype = e.dtype;
  L.key = AKNN_KEY_EUCLIDEAN;
  L.cl = clouds_of(&e);
  L.rows = query_rows(&e);
  L.F = e.F;
  L.K = e.K;
  L.q = q; L.p = p; L.f = f; L.sigma = scale;
  L.idxs = idxs;
  L.dists = dists;  // the scan's Euclidean keys
  L.f_out = f_out;
  L.w_out = w_out;
  hipError_t err;
  {
    Span sp("aggregate_iso", st);
    err = aknn::launch_aggregate_iso(L, st);
  }
  if (err != hipSuccess) return hip_fail(err, "aggregate_iso");
  return AKNN_OK;
}

// Lookup workspace: the 1-NN scan's own workspace, then its key and index per query.
static aknn_dims nn1_dims(const aknn_dims* d) {
  aknn_dims e = *d;
  e.K = 1;
  e.key = AKNN_KEY_EUCLIDEAN;
  return e;
}

size_t aknn_lookup_workspace_size(const aknn_dims* d) {
  if (!d) return 0;
  const aknn_dims e = nn1_dims(d);
  const size_t scan = aknn_forward_workspace_size(&e);
  if (scan == 0) return 0;
  return align256(scan) + align256(static_cast<size_t>(query_rows(d)) * elem_size(d)) +
         align256(static_cast<size_t>(query_rows(d)) * 8);
}

int aknn_lookup_forward(const aknn_dims* d, const void* q, const void* p, const void* f, const void* sigma,
                        const void* R, const int64_t* knn_table, int64_t table_k, void* f_out, void* w_out,
                        void* dists, int64_t* idxs, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (d->q_off) return fail(AKNN_E_UNSUPPORTED, "the table lookup takes dense batches (one [P, K] table)");
  if (table_k < d->K) return fail(AKNN_E_INVALID, "table width %lld < K %lld", (long long)table_k, (long long)d->K);
  if (query_rows(d) == 0 || d->B == 0) return AKNN_OK;
  if (!q || !p || !f || !sigma || !R || !knn_table || !f_out || !w_out || !dists || !idxs)
    return fail(AKNN_E_INVALID, "null tensor pointer");
  const size_t need = aknn_lookup_workspace_size(d);
  if (need == 0) return AKNN_E_LAUNCH;
  if (!ws || ws_bytes < need) return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes, need);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const aknn_dims e = nn1_dims(d);
  const size_t scan = align256(aknn_forward_workspace_size(&e));
  char* base = static_cast<char*>(ws);
  void* d1 = base + scan;
  int64_t* i1 = reinterpret_cast<int64_t*>(base + scan + align256(static_cast<size_t>(query_rows(d)) * elem_size(d)));
  rc = run_knn(&e, nullptr, q, p, sigma, R, d1, i1, ws, scan, st);
  if (rc) return rc;
  hipError_t err;
  {
    Span sp("lookup_gather", st);
    err = aknn::launch_lookup_gather(d1, i1, knn_table, table_k, query_rows(d), d->K, d->dtype, idxs, st);
  }
  if (err != hipSuccess) return hip_fail(err, "lookup_gather");
  aknn::QueryLaunch L{};
  L.dtype = d->dtype;
  L.key = AKNN_KEY_ANISO;
  L.cl = clouds_of(d);
  L.rows = query_rows(d);
  L.F = d->F;
  L.K = d->K;
  L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
  L.idxs = idxs;
  L.dists = nullptr;  // anisotropic D recomputed for the table's neighbours
  L.dists_out = dists;
  L.f_out = f_out;
  L.w_out = w_out;
  {
    Span sp("aggregate", st);
    err = aknn::launch_aggregate(L, st);
  }
  if (err != hipSuccess) return hip_fail(err, "aggregate");
  return AKNN_OK;
}

size_t aknn_csr_workspace_size(const aknn_dims* d) {
  if (check_dims(d, false) != AKNN_OK) return 0;
  size_t bytes = 0;
  const hipError_t e = aknn::csr_temp_bytes(query_rows(d) * d->K, point_rows(d), query_rows(d), &bytes);
  if (e != hipSuccess) {
    hip_fail(e, "csr_workspace_size");
    return 0;
  }
  return bytes;
}

static int build_csr_common(const aknn_dims* d, const int64_t* idxs, const int32_t* qorder, int32_t* csr_offsets,
                            int32_t* csr_entries, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  if (!csr_offsets || (query_rows(d) > 0 && (!idxs || !csr_entries)))
    return fail(AKNN_E_INVALID, "null pointer");
  const size_t need = aknn_csr_workspace_size(d);
  if (need == 0) return AKNN_E_LAUNCH;  // message set by the size query
  if (!ws || ws_bytes < need)
    return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes, need);
  hipError_t e;
  {
    Span sp("build_csr", static_cast<hipStream_t>(stream));
    e = aknn::launch_build_csr(idxs, clouds_of(d), query_rows(d), d->K, point_rows(d), csr_offsets,
                               csr_entries, ws, ws_bytes, static_cast<hipStream_t>(stream), qorder);
  }
  if (e != hipSuccess) return hip_fail(e, "build_csr");
  return AKNN_OK;
}

int aknn_build_csr(const aknn_dims* d, const int64_t* idxs, int32_t* csr_offsets, int32_t* csr_entries, void* ws,
                   size_t ws_bytes, void* stream) {
  return build_csr_common(d, idxs, nullptr, csr_offsets, csr_entries, ws, ws_bytes, stream);
}

int aknn_build_csr_ordered(const aknn_dims* d, const int64_t* idxs, const int32_t* query_order,
                           int32_t* csr_offsets, int32_t* csr_entries, void* ws, size_t ws_bytes, void* stream) {
  if (!query_order && d && query_rows(d) > 0) return fail(AKNN_E_INVALID, "query_order is NULL");
  return build_csr_common(d, idxs, query_order, csr_offsets, csr_entries, ws, ws_bytes, stream);
}

// dq via per-entry contributions when the point kernels run anyway.
bool contrib_path(const aknn_dims* d, uint32_t needs) {
  return (needs & AKNN_NEED_DQ) && ((needs & AKNN_NEED_POINTS) || (needs & AKNN_SECOND_FULL)) &&
         point_rows(d) > 0 && query_rows(d) > 0;
}

struct Upstream2 {  // the second order's upstream beyond grad_grad_q (full second order)
  const void *p, *f, *sigma, *R;
  bool any() const { return p || f || sigma || R; }
};

// Bytes of the point dot products of a broadcast upstream.
size_t dots_bytes(const aknn_dims* d) { return align256(static_cast<size_t>(point_rows(d)) * 16 * elem_size(d)); }

// Workspace of the backward without the ordered query copies.
static size_t backward_base_bytes(const aknn_dims* d, uint32_t needs) {
  if (needs & AKNN_GRAD_BROADCAST) return (needs & AKNN_NEED_DQ) && query_rows(d) > 0 ? dots_bytes(d) : 0;
  if (!contrib_path(d, needs)) return 0;
  return align256(static_cast<size_t>(query_rows(d)) * d->K * 3 * elem_size(d));
}

// Ordered copies of the query coordinates and of grad_grad_q (CSR in the query order).
static size_t ordered_copy_bytes(const aknn_dims* d) {
  return align256(static_cast<size_t>(query_rows(d)) * 3 * elem_size(d));
}

static int backward_common(const aknn_dims* d, uint32_t needs, bool second, const void* gg,
                           const Upstream2& up, const void* g,
                           const void* q, const void* p, const void* f, const void* sigma,
                           const void* R, const int64_t* idxs, const int32_t* csr_off,
                           const int32_t* csr_ent, const aknn_cache* cache_in, void* ws,
                           size_t ws_bytes, void* dgf, void* dq, void* dp, void* df, void* dsigma,
                           void* dR, void* stream, const void* gw = nullptr, void* dgw = nullptr) {
  int rc = check_dims(d, true);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool need_dgw = second && (needs & AKNN_NEED_DGRAD_W_OUT) && query_rows(d) > 0;
  needs &= ~AKNN_NEED_DGRAD_W_OUT;
  const bool csr_ordered = (needs & AKNN_CSR_QUERY_ORDER) != 0;
  needs &= ~AKNN_CSR_QUERY_ORDER;
  if (csr_ordered && query_rows(d) > 0 && !(cache_in && cache_in->query_order))
    return fail(AKNN_E_INVALID, "AKNN_CSR_QUERY_ORDER needs the cache's query order");
  if (need_dgw && !dgw) return fail(AKNN_E_INVALID, "d_grad_w_out requested but NULL");
  const bool full = second && up.any();
  if (full && !(needs & AKNN_SECOND_FULL))
    return fail(AKNN_E_INVALID, "grad_grad_p / _f / _sigma / _R given: set AKNN_SECOND_FULL in needs");
  if (!full) needs &= ~AKNN_SECOND_FULL;
  const bool bcast = (needs & AKNN_GRAD_BROADCAST) != 0;
  if (bcast && full)
    return fail(AKNN_E_INVALID, "the full second order takes a materialised grad_f_out (no AKNN_GRAD_BROADCAST)");
  // A forward over no query rows fills no cache (its orders are left unwritten): ignore it.
  const aknn_cache* cache = query_rows(d) > 0 ? cache_in : nullptr;
  const bool via_contrib = !bcast && contrib_path(d, needs);
  // broadcast upstream: dq from point dot products, unless the query kernel runs for
  // d grad_f_out anyway (it then forms dq alongside)
  const bool dq_dots = bcast && (needs & AKNN_NEED_DQ) && !(second && (needs & AKNN_NEED_DGRAD_F_OUT)) &&
                       query_rows(d) > 0;
  const bool own_dots = dq_dots && !(cache && cache->point_dots);
  if ((via_contrib || own_dots) && (!ws || ws_bytes < aknn_backward_workspace_size(d, needs)))
    return fail(AKNN_E_WORKSPACE, "workspace too small (%zu < %zu bytes)", ws_bytes,
                aknn_backward_workspace_size(d, needs));
  if (!g || !q || !p || !f || !sigma || !R || (query_rows(d) > 0 && !idxs))
    return fail(AKNN_E_INVALID, "null input pointer");
  if ((needs & AKNN_NEED_DQ) && !dq) return fail(AKNN_E_INVALID, "dq requested but NULL");
  if ((needs & AKNN_NEED_DGRAD_F_OUT) && !dgf) return fail(AKNN_E_INVALID, "d_grad_f_out requested but NULL");
  if ((needs & AKNN_NEED_DP) && !dp) return fail(AKNN_E_INVALID, "dp requested but NULL");
  if ((needs & AKNN_NEED_DF) && !df) return fail(AKNN_E_INVALID, "df requested but NULL");
  if ((needs & AKNN_NEED_DSIGMA) && !dsigma) return fail(AKNN_E_INVALID, "dsigma requested but NULL");
  if ((needs & AKNN_NEED_DR) && !dR) return fail(AKNN_E_INVALID, "dR requested but NULL");
  if (dq_dots) {
    const void* dots = own_dots ? ws : cache->point_dots;
    hipError_t e;
    if (own_dots) {
      aknn::PointLaunch D{};
      D.dtype = d->dtype;
      D.rows = point_rows(d);
      D.F = d->F;
      D.K = d->K;
      D.f = f;
      D.p = p;
      D.sigma = sigma;
      D.R = R;
      D.grad_f_out = g;
      D.df = (needs & AKNN_NEED_DF) ? df : nullptr;  // same vector width as the point kernel
      D.g_bcast = true;
      D.cl = clouds_of(d);
      {
        Span sp("point_dots", st);
        e = aknn::launch_point_dots(D, ws, st);
      }
      if (e != hipSuccess) return hip_fail(e, "point_dots");
    }
    aknn::QueryLaunch L{};
    L.dtype = d->dtype;
    L.key = d->key;
    L.cl = clouds_of(d);
    L.rows = query_rows(d);
    L.F = d->F;
    L.K = d->K;
    L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
    L.idxs = idxs;
    L.grad_grad_q = gg;
    L.dq = dq;
    L.order = cache ? cache->query_order : nullptr;
    L.rec = cache ? cache->point_records : nullptr;
    L.dots = dots;
    L.gw = gw;
    L.dgw = need_dgw ? dgw : nullptr;
    {
      Span sp(second ? "dq_dots2" : "dq_dots", st);
      e = aknn::launch_dq_dots(L, second, st);
    }
    if (e != hipSuccess) return hip_fail(e, "dq_dots");
  }
  const bool dgw_by_dots = dq_dots && second;  // the dot-record kernel wrote d grad_w_out
  const bool want_q = ((needs & AKNN_NEED_DQ) && !via_contrib && !dq_dots) ||
                      (second && (needs & AKNN_NEED_DGRAD_F_OUT)) || (need_dgw && !dgw_by_dots);
  if (want_q && query_rows(d) > 0) {
    aknn::QueryLaunch L{};
    L.dtype = d->dtype;
    L.key = d->key;
    L.cl = clouds_of(d);
    L.rows = query_rows(d);
    L.F = d->F;
    L.K = d->K;
    L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
    L.idxs = idxs;
    L.grad_f_out = g;
    L.grad_grad_q = gg;
    L.gg_p = up.p;
    L.gg_sigma = up.sigma;
    L.gg_R = up.R;
    L.dq = ((needs & AKNN_NEED_DQ) && !via_contrib) ? dq : nullptr;
    L.d_grad_f_out = (needs & AKNN_NEED_DGRAD_F_OUT) ? dgf : nullptr;
    L.order = cache ? cache->query_order : nullptr;
    L.rec = cache ? cache->point_records : nullptr;
    L.g_bcast = bcast;
    L.gw = gw;
    L.dgw = (need_dgw && !dgw_by_dots) ? dgw : nullptr;
    hipError_t e;
    {
      Span sp(full ? "query_grad2_full" : (second ? "query_grad2" : "query_grad"), st);
      e = full ? aknn::launch_query_grad_full(L, st) : aknn::launch_query_grad(L, second, st);
    }
    if (e != hipSuccess) return hip_fail(e, second ? "query_grad2" : "query_grad");
    if (full && up.f && (needs & AKNN_NEED_DGRAD_F_OUT)) {
      // d grad_f_out += sum_k w_k grad_grad_f[n_k]: the forward aggregation of gg_f
      aknn::QueryLaunch A = L;
      A.f = up.f;
      A.dists = nullptr;
      A.dists_out = nullptr;
      A.w_out = nullptr;
      A.f_out = dgf;
      A.accumulate = true;
      A.key = AKNN_KEY_ANISO;
      {
        Span sp("query_grad2_ggf", st);
        e = aknn::launch_aggregate(A, st);
      }
      if (e != hipSuccess) return hip_fail(e, "query_grad2_ggf");
    }
  }
  if (((needs & AKNN_NEED_POINTS) || via_contrib) && point_rows(d) > 0) {
    if (!csr_off || (query_rows(d) > 0 && !csr_ent))
      return fail(AKNN_E_INVALID, "point gradients need the CSR from aknn_build_csr");
    // CSR in the spatial query order: the query-side rows the point kernels gather, in
    // that order, after the workspace's other contents
    void *q_sorted = nullptr, *hh_sorted = nullptr;
    if (csr_ordered && query_rows(d) > 0) {
      const size_t base = backward_base_bytes(d, needs | (full ? AKNN_SECOND_FULL : 0u));
      const size_t total = base + 2 * ordered_copy_bytes(d);
      if (!ws || ws_bytes < total)
        return fail(AKNN_E_WORKSPACE, "workspace too small for AKNN_CSR_QUERY_ORDER (%zu < %zu bytes)", ws_bytes,
                    total);
      q_sorted = static_cast<char*>(ws) + base;
      hh_sorted = static_cast<char*>(q_sorted) + ordered_copy_bytes(d);
      hipError_t e = aknn::launch_gather3(q, cache->query_order, q_sorted, query_rows(d), d->dtype, st);
      if (e == hipSuccess && gg) e = aknn::launch_gather3(gg, cache->query_order, hh_sorted, query_rows(d), d->dtype, st);
      if (e != hipSuccess) return hip_fail(e, "query_order copies");
    }
    aknn::PointLaunch L{};
    L.dtype = d->dtype;
    L.key = d->key;
    L.rows = point_rows(d);
    L.F = d->F;
    L.K = d->K;
    L.csr_off = csr_off;
    L.csr_ent = csr_ent;
    L.q = q; L.p = p; L.f = f; L.sigma = sigma; L.R = R;
    L.grad_f_out = g;
    L.second = second;
    L.grad_grad_q = gg;
    L.gg_p = up.p;
    L.gg_f = up.f;
    L.gg_sigma = up.sigma;
    L.gg_R = up.R;
    L.dp = (needs & AKNN_NEED_DP) ? dp : nullptr;
    L.df = (needs & AKNN_NEED_DF) ? df : nullptr;
    L.dsigma = (needs & AKNN_NEED_DSIGMA) ? dsigma : nullptr;
    L.dR = (needs & AKNN_NEED_DR) ? dR : nullptr;
    L.order = cache ? cache->point_order : nullptr;
    L.dq_contrib = via_contrib ? ws : nullptr;
    L.g_bcast = bcast;
    L.cl = clouds_of(d);
    L.gw = gw;
    if (q_sorted) {
      L.qorder = cache->query_order;
      L.q_sorted = q_sorted;
      L.hh_sorted = gg ? hh_sorted : nullptr;
    }
    hipError_t e;
    {
      Span sp(full ? "point_grad2_full" : (second ? "point_grad2" : "point_grad"), st);
      e = aknn::launch_point_grad(L, st);
    }
    if (e != hipSuccess) return hip_fail(e, second ? "point_grad2" : "point_grad");
    if (via_contrib) {
      {
        Span sp(second ? "dq_reduce2" : "dq_reduce", st);
        e = aknn::launch_contrib_reduce(ws, query_rows(d), d->K, dq, d->dtype, st);
      }
      if (e != hipSuccess) return hip_fail(e, "dq_reduce");
    }
  }
  return AKNN_OK;
}

size_t aknn_backward_workspace_size(const aknn_dims* d, uint32_t needs) {
  if (check_dims(d, false) != AKNN_OK) return 0;
  size_t bytes = backward_base_bytes(d, needs);
  if (needs & AKNN_CSR_QUERY_ORDER) bytes += 2 * ordered_copy_bytes(d);
  return bytes;
}

int aknn_point_dots(const aknn_dims* d, const void* grad_rows, const void* f, const void* p, const void* sigma,
                    const void* R, void* a, void* stream) {
  int rc = check_dims(d, false);
  if (rc) return rc;
  if (point_rows(d) == 0) return AKNN_OK;
  if (!grad_rows || !f || !p || !sigma || !R || !a) return fail(AKNN_E_INVALID, "null pointer");
  aknn::PointLaunch D{};
  D.dtype = d->dtype;
  D.rows = point_rows(d);
  D.F = d->F;
  D.K = d->K;
  D.f = f;
  D.p = p;
  D.sigma = sigma;
  D.R = R;
  D.grad_f_out = grad_rows;
  D.g_bcast = true;
  D.cl = clouds_of(d);
  hipError_t e;
  {
    Span sp("point_dots", static_cast<hipStream_t>(stream));
    e = aknn::launch_point_dots(D, a, static_cast<hipStream_t>(stream));
  }
  if (e != hipSuccess) return hip_fail(e, "point_dots");
  return AKNN_OK;
}

int aknn_backward_w(const aknn_dims* d, uint32_t needs, const void* grad_f_out, const void* grad_w_out,
                    const void* q, const void* p, const void* f, const void* sigma, const void* R,
                    const int64_t* idxs, const int32_t* csr_offsets, const int32_t* csr_entries,
                    const aknn_cache* cache, void* ws, size_t ws_bytes, void* dq, void* dp, void* df,
                    void* dsigma, void* dR, void* stream) {
  return backward_common(d, needs & ~(AKNN_NEED_DGRAD_F_OUT | AKNN_SECOND_FULL | AKNN_NEED_DGRAD_W_OUT), false,
                         nullptr, Upstream2{}, grad_f_out, q, p, f, sigma, R, idxs, csr_offsets, csr_entries,
                         cache, ws, ws_bytes, nullptr, dq, dp, df, dsigma, dR, stream, grad_w_out, nullptr);
}

int aknn_backward2_w(const aknn_dims* d, uint32_t needs, const void* grad_grad_q, const void* grad_grad_p,
                     const void* grad_grad_f, const void* grad_grad_sigma, const void* grad_grad_R,
                     const void* grad_f_out, const void* grad_w_out, const void* q, const void* p,
                     const void* f, const void* sigma, const void* R, const int64_t* idxs,
                     const int32_t* csr_offsets, const int32_t* csr_entries, const aknn_cache* cache, void* ws,
                     size_t ws_bytes, void* d_grad_f_out, void* d_grad_w_out, void* dq, void* dp, void* df,
                     void* dsigma, void* dR, void* stream) {
  const Upstream2 up{grad_grad_p, grad_grad_f, grad_grad_sigma, grad_grad_R};
  if (!grad_grad_q && !up.any()) return fail(AKNN_E_INVALID, "every upstream gradient is NULL");
  const uint32_t mode = up.any() ? (needs | AKNN_SECOND_FULL) : (needs & ~AKNN_SECOND_FULL);
  return backward_common(d, mode, true, grad_grad_q, up, grad_f_out, q, p, f, sigma, R, idxs, csr_offsets,
                         csr_entries, cache, ws, ws_bytes, d_grad_f_out, dq, dp, df, dsigma, dR, stream,
                         grad_w_out, d_grad_w_out);
}

int aknn_backward(const aknn_dims* d, uint32_t needs, const void* grad_f_out, const void* q,
                  const void* p, const void* f, const void* sigma, const void* R,
                  const int64_t* idxs, const int32_t* csr_offsets, const int32_t* csr_entries,
                  const aknn_cache* cache, void* ws, size_t ws_bytes, void* dq, void* dp, void* df,
                  void* dsigma, void* dR, void* stream) {
  return backward_common(d, needs & ~(AKNN_NEED_DGRAD_F_OUT | AKNN_SECOND_FULL | AKNN_NEED_DGRAD_W_OUT), false,
                         nullptr, Upstream2{},
                         grad_f_out, q, p, f, sigma, R,
                         idxs, csr_offsets, csr_entries, cache, ws, ws_bytes, nullptr, dq, dp, df,
                         dsigma, dR, stream);
}

int aknn_backward2(const aknn_dims* d, uint32_t needs, const void* grad_grad_q,
                   const void* grad_f_out, const void* q, const void* p, const void* f,
                   const void* sigma, const void* R, const int64_t* idxs,
                   const int32_t* csr_offsets, const int32_t* csr_entries,
                   const aknn_cache* cache, void* ws, size_t ws_bytes, void* d_grad_f_out,
                   void* dq, void* dp, void* df, void* dsigma, void* dR, void* stream) {
  if (!grad_grad_q) return fail(AKNN_E_INVALID, "grad_grad_q is NULL");
  return backward_common(d, needs & ~(AKNN_SECOND_FULL | AKNN_NEED_DGRAD_W_OUT), true, grad_grad_q, Upstream2{},
                         grad_f_out, q, p, f,
                         sigma, R, idxs, csr_offsets, csr_entries, cache, ws, ws_bytes, d_grad_f_out, dq, dp,
                         df, dsigma, dR, stream);
}

int aknn_backward2_full(const aknn_dims* d, uint32_t needs, const void* grad_grad_q, const void* grad_grad_p,
                        const void* grad_grad_f, const void* grad_grad_sigma, const void* grad_grad_R,
                        const void* grad_f_out, const void* q, const void* p, const void* f,
                        const void* sigma, const void* R, const int64_t* idxs,
                        const int32_t* csr_offsets, const int32_t* csr_entries,
                        const aknn_cache* cache, void* ws, size_t ws_bytes, void* d_grad_f_out,
                        void* dq, void* dp, void* df, void* dsigma, void* dR, void* stream) {
  const Upstream2 up{grad_grad_p, grad_grad_f, grad_grad_sigma, grad_grad_R};
  if (!grad_grad_q && !up.any()) return fail(AKNN_E_INVALID, "every upstream gradient is NULL");
  return backward_common(d, (needs | AKNN_SECOND_FULL) & ~AKNN_NEED_DGRAD_W_OUT, true, grad_grad_q, up,
                         grad_f_out, q, p, f, sigma, R,
                         idxs, csr_offsets, csr_entries, cache, ws, ws_bytes, d_grad_f_out, dq, dp, df, dsigma,
                         dR, stream);
}

}  // extern "C"
