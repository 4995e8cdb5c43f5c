// Exact top-K with spatial tile pruning (gfx950).
//
// Same result as the brute-force scan (csrc/forward.cu:83-97 semantics, (key, index)
// order) in far fewer key evaluations.  Per call:
//
//   1. cloud boxes -> 24-bit Morton code per point and per query (8 bits per axis in
//      the cloud's point box); stable radix sorts give a spatial order of both.
//   2. Points are regrouped in that order into tiles of kTile records
//      {p0 p1 p2, R00..R22, s0 s1 s2, original index}; each tile keeps its box and
//      mu_tile = min over its points of a certified lower bound mu_n such that the
//      FP32 key the scan computes satisfies   key(q, n) >= mu_n * |q - p_n|^2.
//   3. One wave = 64 spatially sorted queries.  It visits the cloud's tiles outward
//      from the tile nearest its first query; a tile is scanned only if some lane has
//      mu_tile * dist^2(q, box) <= its current K-th key, and within it only the
//      8-record sub-tiles whose own box and mu pass the same test.  Skipped points
//      provably have keys above every lane's K-th, so the kept set and order are those
//      of a full scan.
//
// mu_n (point with M = R^T diag(s) R): for the exact D = d^T M d >= lmin(M) |d|^2, and
// for the scan's FP32 evaluation (one rounding per op, unit roundoff u, g = 4u/(1-4u)):
//   key >= (1 - g) (D - 2 g |d| sqrt(D S)),   S = sum_i s_i |R_i|^2 = trace(M)
// which is increasing in D where it matters, hence key >= (1-g)(lmin - 2 g sqrt(lmin S)) |d|^2.
// lmin comes from Jacobi rotations of the 3x3 matrix in FP64 minus a margin, the
// result is scaled by (1 - 1e-6) to absorb the FP32 rounding of the tile test itself.
// Points with any negative scale, or non-finite data, mark their tile "always scan".
#include <map>
#include <mutex>
#include <utility>

#include "aknn_bounds.h"
#include "aknn_common.h"
#include "aknn_internal.h"
#include "aknn_order.h"

namespace aknn {

constexpr int kSubsMax = kTile / kSubRecMin;  // sub-tile headers kept per tile
constexpr int kTileSentinel = kNoPoint;
constexpr int kWaveSerialInserts = 16;  // per-query walk: more candidates in a tile -> sort

// Workspace pieces of the tiled search (see carve()).
struct PrunedState {
  float* box;
  unsigned* obox;
  uint64_t *pkeys, *pkeys_s, *qkeys, *qkeys_s;
  uint32_t *pperm, *qperm;
  int64_t *ptp, *qtp, *stp;
  void* rec;
  float *hdr, *shdr, *sub;  // tile, super-tile and sub-tile headers ([tiles][kSubsMax][8])
  int32_t* tcnt;
  void* tmp;
  size_t tmp_bytes;
  void* lkey;       // split walk: per (sorted query, way) top-K keys   [qrows][ways][K]
  int32_t* lid;     //                                    and indices
  void* bnd;        // split walk: shared per-query bound (ordered encoding) [qrows]
  unsigned* work;   // persistent search: next-item counters, one per XCD [kXcds]
  int32_t* torder;  // walk items: (query tile, part) in walk order, widest first per XCD range [walk_items]
  size_t total;
};

// Split walks: per kSplitPer query tiles of each XCD's range, the kSplitTiles widest are
// walked as kSplitParts parts of 64 / kSplitParts queries each (one wave per part; the
// other lanes hold no query), so kSplitTiles * (kSplitParts - 1) extra walk items.  A wide tile's 64 queries straddle far-apart regions;
// a part's 16 lie closer together, so each part walks a fraction of the tile's union
// of neighbourhoods and the longest walks, which otherwise set the kernel's end, are
// shortened.  Walk item codes: tile << 3 | part << 1 | split.
constexpr int64_t kSplitPer = 128, kSplitTiles = 1, kSplitParts = 4;
__host__ __device__ inline int64_t split_extra(int64_t ntiles) {  // extra items per XCD range
  return ntiles < (int64_t(1) << 27) ? ntiles / (kSplitPer * kXcds) * kSplitTiles * (kSplitParts - 1) : 0;
}
__host__ __device__ inline int64_t walk_items(int64_t ntiles) { return ntiles + kXcds * split_extra(ntiles); }

// Order-preserving unsigned encodings of the key (for atomic min on the shared bound).
template <typename T>
struct OrdKey;
template <>
struct OrdKey<float> {
  using U = unsigned;
  static __device__ __forceinline__ U enc(float f) {
    const U u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
  static __device__ __forceinline__ float dec(U k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
  }
};
template <>
struct OrdKey<double> {
  using U = unsigned long long;
  static __device__ __forceinline__ U enc(double f) {
    const U u = static_cast<U>(__double_as_longlong(f));
    return (u >> 63) ? ~u : (u | (1ull << 63));
  }
  static __device__ __forceinline__ double dec(U k) {
    return __longlong_as_double(static_cast<long long>((k >> 63) ? (k & ~(1ull << 63)) : ~k));
  }
};

// ------------------------------------------------------------------ Morton codes
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 24-bit Morton code (8 bits per axis) in the cloud's box.
constexpr int kCodeBits = 24;
template <typename T>
__device__ __forceinline__ uint32_t morton(T x, T y, T z, const float* box) {
  auto q = [](T v, float lo, float hi) -> uint32_t {
    const float ext = hi - lo;
    float t = ext > 0.f ? (static_cast<float>(v) - lo) / ext : 0.f;
    t = t != t ? 0.f : fminf(fmaxf(t, 0.f), 1.f);  // NaN -> 0; clamp queries outside the box
    return static_cast<uint32_t>(t * 255.f);
  };
  return (spread10(q(x, box[0], box[3])) << 2) | (spread10(q(y, box[1], box[4])) << 1) |
         spread10(q(z, box[2], box[5]));
}

// Cloud boxes: box[b] = {lo x y z, hi x y z, 0, 0} of the cloud's finite points (float).
// Many blocks per cloud reduce into order-preserving integer encodings with atomic
// min/max (exact, so the result does not depend on timing), then a finalize pass
// decodes them.
__device__ __forceinline__ unsigned ord_enc(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void box_init_kernel(unsigned* __restrict__ ob, int64_t B) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * 6) return;
  ob[i] = (i % 6) < 3 ? ord_enc(__builtin_inff()) : ord_enc(-__builtin_inff());
}

constexpr int kBoxChunk = 4096;  // points per reducing block

template <typename T>
__global__ __launch_bounds__(256) void cloud_box_kernel(const T* __restrict__ p, Clouds cl,
                                                        unsigned* __restrict__ ob) {
  const int64_t b = blockIdx.y;
  const int64_t s = cl.p_begin(b) + static_cast<int64_t>(blockIdx.x) * kBoxChunk;
  const int64_t e = min(cl.p_end(b), s + kBoxChunk);
  if (s >= e) return;
  float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  for (int64_t i = s + threadIdx.x; i < e; i += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = static_cast<float>(p[i * 3 + c]);
      if (__builtin_isfinite(v)) {
        lo[c] = fminf(lo[c], v);
        hi[c] = fmaxf(hi[c], v);
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lo[c] = fminf(lo[c], __shfl_xor(lo[c], off, 64));
      hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], off, 64));
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      atomicMin(ob + b * 6 + c, ord_enc(lo[c]));
      atomicMax(ob + b * 6 + 3 + c, ord_enc(hi[c]));
    }
  }
}

__global__ void box_final_kernel(const unsigned* __restrict__ ob, int64_t B, float* __restrict__ box) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * 8) return;
  const int64_t b = i / 8;
  const int c = static_cast<int>(i % 8);
  float v = 0.f;
  if (c < 6) {
    v = ord_dec(ob[b * 6 + c]);
    if (!__builtin_isfinite(v)) v = 0.f;  // cloud without finite points
  }
  box[i] = v;
}

// keys[i] = (cloud << 24) | morton(row i); rows are query rows (QUERY) or point rows.
template <typename T, bool QUERY>
__global__ void codes_kernel(const T* __restrict__ xyz, Clouds cl, int64_t rows,
                             const float* __restrict__ box, uint64_t* __restrict__ keys) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int64_t b = QUERY ? cl.cloud_of_query(i) : cl.cloud_of_point(i);
  const uint32_t c = morton<T>(xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2], box + b * 8);
  keys[i] = (static_cast<uint64_t>(b) << kCodeBits) | c;
}

// ------------------------------------------------------------------ point bound
// point_mu, stored_mu and box_needed: aknn_bounds.h (shared with the host test of the
// certificate).

// ------------------------------------------------------------------ tile build
// One wave per tile.  rec: [tiles*kTile][16] records; hdr: [tiles][8] =
// {lo x y z, hi x y z, mu (float, rounded down; < 0 = always scan), first code (bits)};
// sub: [tiles][kSubsMax][8], of which the first kTile / SUBREC hold the same box and mu
// (last value unused) of each SUBREC-record sub-tile.
//
// Record order within a tile (mu_classes = m): the tile's points are ranked by their
// bound mu ("always scan" points first) and cut into m classes of equal size; records
// are stored class by class, in Morton order within a class.  m = 1 keeps the Morton
// order (sub-tiles are the tightest boxes); larger m groups the points that a weak
// bound lets far queries reach (small mu: thin Gaussians) into their own sub-tiles, so
// the others are skipped from further away.  The set of records, and hence every
// result, does not depend on the order.
template <typename T, int KEY, int SUBREC>
__global__ __launch_bounds__(64) void tile_build_kernel(
    const T* __restrict__ p, const T* __restrict__ sigma, const T* __restrict__ R, Clouds cl,
    const int64_t* __restrict__ tile_prefix, const uint32_t* __restrict__ perm,
    const uint64_t* __restrict__ sorted_keys, T* __restrict__ rec, float* __restrict__ hdr,
    float* __restrict__ sub, int32_t* __restrict__ tile_cnt, int mu_classes) {
  const int64_t t = blockIdx.x;
  const int lane = threadIdx.x;
  int64_t b = 0;
  if (cl.p_off) {
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (tile_prefix[mid] <= t) lo = mid; else hi = mid - 1;
    }
    b = lo;
  } else {
    b = t / tile_prefix[1];  // dense: tile_prefix[1] = tiles per cloud
  }
  const int64_t lt = t - (cl.p_off ? tile_prefix[b] : b * tile_prefix[1]);
  const int64_t pb = cl.p_begin(b);
  const int64_t np = cl.p_end(b) - pb;
  const int64_t i0 = pb + lt * kTile;
  const int cnt = static_cast<int>(min<int64_t>(kTile, np - lt * kTile));
  float lo3[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float hi3[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  double mu = __builtin_inf();
  bool bad = false;
  T v[16];
  if (lane < cnt) {
    const int64_t row = perm[i0 + lane];
    for (int c = 0; c < 3; ++c) v[c] = p[row * 3 + c];
    // R / sigma may be absent for the Euclidean key, which never reads them
    for (int c = 0; c < 9; ++c) v[3 + c] = R ? R[row * 9 + c] : T(0);
    for (int c = 0; c < 3; ++c) v[12 + c] = sigma ? sigma[row * 3 + c] : T(0);
    const int idx = static_cast<int>(row - pb);
    if constexpr (sizeof(T) == 4) v[15] = __int_as_float(idx);
    else v[15] = __longlong_as_double(static_cast<long long>(idx));
    for (int c = 0; c < 3; ++c) {
      // box in float, widened outward so it contains the exact coordinate
      const float f = static_cast<float>(v[c]);
      if (!__builtin_isfinite(f)) bad = true;
      lo3[c] = static_cast<double>(f) <= static_cast<double>(v[c]) ? f : next_down(f);
      hi3[c] = static_cast<double>(f) >= static_cast<double>(v[c]) ? f : next_up(f);
    }
    mu = point_mu<T, KEY == AKNN_KEY_EUCLIDEAN>(v + 3, v + 12);
    if (!(mu >= 0.0)) bad = true;
  } else {
    // padding record: NaN position, so every key evaluated against it is NaN and it
    // never enters a top-K list (the scan reads whole sub-tiles)
    for (int c = 0; c < 16; ++c) v[c] = c < 3 ? static_cast<T>(__builtin_nan("")) : T(0);
    if constexpr (sizeof(T) == 4) v[15] = __int_as_float(kTileSentinel);
    else v[15] = __longlong_as_double(static_cast<long long>(kTileSentinel));
  }
  // slot of this lane's record in the tile (a permutation of the lanes; padding last)
  int slot = lane;
  const int classes = mu_classes;
  const double mkey = lane < cnt ? (bad ? -1.0 : mu) : __builtin_inf();
  if (classes > 1) {
    int r = 0;  // rank by (mu, lane)
    for (int j = 0; j < 64; ++j) {
      const double kj = __shfl(mkey, j, 64);
      r += (kj < mkey) | ((kj == mkey) & (j < lane));
    }
    const int cls = lane < cnt ? r * classes / cnt : classes;
    slot = 0;  // rank by (class, lane)
    for (int j = 0; j < 64; ++j) {
      const int cj = __shfl(cls, j, 64);
      slot += (cj < cls) | ((cj == cls) & (j < lane));
    }
  }
  T* out = rec + (t * kTile + slot) * 16;
  for (int c = 0; c < 16; ++c) out[c] = v[c];
  if (classes > 1) {
    // the box values move to lane `slot` (a forward permute: every lane receives one)
    auto push = [&](int x) { return __builtin_amdgcn_ds_permute(slot << 2, x); };
    for (int c = 0; c < 3; ++c) {
      lo3[c] = __int_as_float(push(__float_as_int(lo3[c])));
      hi3[c] = __int_as_float(push(__float_as_int(hi3[c])));
    }
    const long long mb = __double_as_longlong(mu);
    const long long lo_ = static_cast<uint32_t>(push(static_cast<int>(mb & 0xffffffffll)));
    const long long hi_ = static_cast<long long>(push(static_cast<int>(mb >> 32))) << 32;
    mu = __longlong_as_double(hi_ | lo_);
    bad = push(static_cast<int>(bad)) != 0;
  }
  // Reductions over groups of SUBREC lanes (the sub-tiles), then over the wave (the tile).
  // min / max are exact, so the grouping does not change any value.
  auto write_box = [&](float* h) {
    for (int c = 0; c < 3; ++c) {
      h[c] = lo3[c];
      h[3 + c] = hi3[c];
    }
    h[6] = stored_mu(mu, bad);
  };
  for (int off = 1; off < 64; off <<= 1) {
    for (int c = 0; c < 3; ++c) {
      lo3[c] = fminf(lo3[c], __shfl_xor(lo3[c], off, 64));
      hi3[c] = fmaxf(hi3[c], __shfl_xor(hi3[c], off, 64));
    }
    mu = fmin(mu, __shfl_xor(mu, off, 64));
    bad = bad | (__shfl_xor(static_cast<int>(bad), off, 64) != 0);
    if (off == SUBREC / 2 && (lane % SUBREC) == 0) {
      float* h = sub + (t * kSubsMax + lane / SUBREC) * 8;
      write_box(h);
      h[7] = 0.f;
    }
  }
  if (lane == 0) {
    float* h = hdr + t * 8;
    write_box(h);
    h[7] = __uint_as_float(static_cast<uint32_t>(sorted_keys[i0] & 0xffffffffu));
    tile_cnt[t] = cnt;
  }
}

// ------------------------------------------------------------------ pruned top-K
template <int N, typename T>
__device__ __forceinline__ void lex_insert(T (&k)[N], int (&v)[N], T c, int ci) {
  bool lt[N];
#pragma unroll
  for (int i = 0; i < N; ++i) lt[i] = (c < k[i]) | ((c == k[i]) & (ci < v[i]));
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

// Same result as lex_insert, as a chain of compare-and-swap steps: only one compare
// mask is live at a time, which keeps the scalar registers free for the records.
template <int N, typename T>
__device__ __forceinline__ void bubble_insert(T (&k)[N], int (&v)[N], T c, int ci) {
  T ck = c;
  int cv = ci;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const bool sw = (ck < k[i]) | ((ck == k[i]) & (cv < v[i]));
    const T tk = k[i];
    const int tv = v[i];
    k[i] = sw ? ck : tk;
    v[i] = sw ? cv : tv;
    ck = sw ? tk : ck;
    cv = sw ? tv : cv;
  }
}

// A lane's top-K list of the tiled search, ascending in (key, index) order.
// Generic form: separate key and index arrays, inserted by bubble_insert.
template <typename T, int N>
struct TopList {
  T k[N];
  int v[N];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      k[s] = __builtin_inf();
      v[s] = kNoPoint;
    }
  }
  __device__ __forceinline__ T last() const { return k[N - 1]; }
  __device__ __forceinline__ void insert(T c, int ci) { bubble_insert<N>(k, v, c, ci); }
  __device__ __forceinline__ T key(int s) const { return k[s]; }
  __device__ __forceinline__ int id(int s) const { return v[s]; }
};

// FP64 min / max of two list words (aknn_order.h: finite doubles, never NaN), written
// out so that the compiler adds no NaN-quieting canonicalisation of their operands.
__device__ __forceinline__ double word_max(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double word_min(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// FP32 form: one word per entry, (key, index) encoded as a finite double whose numeric
// order is (key, index) order (aknn_order.h), so that an insertion is
// w[i] = min(w[i], max(w[i-1], x)) on the old words: two FP64 min/max per slot, in
// place.  -0 is encoded as +0 (they compare equal as keys); a -0 key is restored on
// output by the caller.  NaN keys are never inserted.  `lastk` caches the decoded last
// key for the pruning tests.
template <int N>
struct TopList<float, N> {
  double w[N];
  float lastk;
  static __device__ __forceinline__ double word(float c, int ci) {
    return __longlong_as_double(static_cast<long long>(list_word(c, ci)));
  }
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int s = 0; s < N; ++s) w[s] = word(__builtin_inff(), kNoPoint);
    lastk = __builtin_inff();
  }
  __device__ __forceinline__ float last() const { return lastk; }
  __device__ __forceinline__ void insert(float c, int ci) {
    const double x = word(c, ci);
    // slots from the last down, two at a time (independent pairs, so that no result is
    // read by the next instruction); each reads w[i-1] before it is updated
#pragma unroll
    for (int i = N - 1; i >= 2; i -= 2) {
      const double ta = word_max(w[i - 1], x);
      const double tb = word_max(w[i - 2], x);
      w[i] = word_min(w[i], ta);
      w[i - 1] = word_min(w[i - 1], tb);
    }
    // left: slots 1 and 0 (N even) or slot 0 (N odd)
    if constexpr (N % 2 == 0) w[1] = word_min(w[1], word_max(w[0], x));
    w[0] = word_min(w[0], x);
    lastk = key(N - 1);
  }
  __device__ __forceinline__ float key(int s) const {
    return list_word_key(static_cast<uint64_t>(__double_as_longlong(w[s])));
  }
  __device__ __forceinline__ int id(int s) const {
    return list_word_index(static_cast<uint64_t>(__double_as_longlong(w[s])));
  }
};

// One 16-value point record, read with wave-uniform addresses (scalar loads).
template <typename T>
struct Rec16 {
  T v[16];
};
template <typename T>
__device__ __forceinline__ Rec16<T> load_rec16(const T* __restrict__ r) {
  Rec16<T> o;
  if constexpr (sizeof(T) == 4) {
    const float4* r4 = reinterpret_cast<const float4*>(r);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 x = r4[i];
      o.v[4 * i] = x.x;
      o.v[4 * i + 1] = x.y;
      o.v[4 * i + 2] = x.z;
      o.v[4 * i + 3] = x.w;
    }
  } else {
    const double2* r2 = reinterpret_cast<const double2*>(r);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double2 x = r2[i];
      o.v[2 * i] = x.x;
      o.v[2 * i + 1] = x.y;
    }
  }
  return o;
}
template <typename T, int KEY>
__device__ __forceinline__ T rec_key(T qx, T qy, T qz, const Rec16<T>& r) {
  return pair_key<T, KEY>(qx, qy, qz, r.v[0], r.v[1], r.v[2], r.v[3], r.v[4], r.v[5], r.v[6], r.v[7],
                          r.v[8], r.v[9], r.v[10], r.v[11], r.v[12], r.v[13], r.v[14]);
}

template <typename T>
__device__ __forceinline__ int rec_index(T v) {
  if constexpr (sizeof(T) == 4) return __float_as_int(v);
  else return static_cast<int>(__double_as_longlong(v));
}

// A box header {lo x y z, hi x y z, mu, -} read with a wave-uniform address (scalar
// loads), and whether any lane still needs the box: mu * dist^2(q, box) not above the
// lane's K-th key thr (box_needed; lanes without a query hold thr = -inf), decided on the
// wave's ballot.  A negative mu ("always scan") is tested on the scalar unit: such a box
// is needed by the wave (its first lane is always a query).
struct BoxHdr {
  float v[8];
};
__device__ __forceinline__ BoxHdr load_hdr(const float* __restrict__ h) {
  BoxHdr o;
  const float4* h4 = reinterpret_cast<const float4*>(h);
  const float4 a = h4[0], b = h4[1];
  o.v[0] = a.x; o.v[1] = a.y; o.v[2] = a.z; o.v[3] = a.w;
  o.v[4] = b.x; o.v[5] = b.y; o.v[6] = b.z; o.v[7] = b.w;
  return o;
}
template <typename T>
__device__ __forceinline__ bool box_any(const BoxHdr& h, T qx, T qy, T qz, T thr) {
  // branch-free, so that the loads of a batch of headers stay together
  const bool always = __float_as_int(h.v[6]) < 0;
  const T lb = box_lower_bound<T>(h.v[0], h.v[1], h.v[2], h.v[3], h.v[4], h.v[5], h.v[6], qx, qy, qz);
  return always | (__builtin_amdgcn_ballot_w64(!(lb > thr)) != 0);
}

// Outward walk over [lo, hi) from c: c, c+1, c-1, c+2, c-2, ... (each index once).
struct Outward {
  int up, down, lo, hi;
  __device__ Outward(int lo_, int hi_, int c) : up(c), down(c - 1), lo(lo_), hi(hi_) {}
  __device__ __forceinline__ int next(int step) {
    const bool take_up = (up < hi) && ((step & 1) == 0 || down < lo);
    return take_up ? up++ : down--;
  }
};

// Occupancy target of the tiled search: its list (KMAX keys and indices, with the copy
// the insertion chain keeps) must fit in the VGPR budget, or it spills to scratch.
template <typename T, int KMAX>
constexpr int search_waves() {
  if constexpr (sizeof(T) == 4) return KMAX <= 16 ? 6 : KMAX <= 32 ? 3 : 2;
  else return KMAX <= 8 ? 6 : KMAX <= 16 ? 4 : KMAX <= 32 ? 2 : 1;
}

// grid.x = query tiles of 64 sorted queries (ragged: via the query tile prefix),
// grid.y = clouds (dense).  Two-level walk: super-tiles of `super` tiles, then tiles,
// then SUBREC-record sub-tiles.  WAVES: occupancy target (0 = the list's default).
template <typename T, int KMAX, int KEY, int SUBREC, int WAVES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES ? WAVES : search_waves<T, KMAX>(),
                                                                   WAVES ? WAVES : search_waves<T, KMAX>()))) void knn_pruned_kernel(
    const T* __restrict__ q, Clouds cl, const int64_t* __restrict__ qtiles,
    const uint32_t* __restrict__ qperm, const uint64_t* __restrict__ qkeys, const int32_t* __restrict__ torder,
    const int64_t* __restrict__ ptiles, const int64_t* __restrict__ stiles,
    const T* __restrict__ rec, const float* __restrict__ hdr, const float* __restrict__ shdr,
    const float* __restrict__ sub, const int32_t* __restrict__ tcnt, int K, T* __restrict__ dists,
    int64_t* __restrict__ idxs, unsigned long long* __restrict__ stats, int ways, T* __restrict__ lkey,
    int32_t* __restrict__ lid, typename OrdKey<T>::U* __restrict__ bnd, const T* __restrict__ pts,
    const T* __restrict__ sig, const T* __restrict__ rot, int super, unsigned* __restrict__ work,
    int64_t nitems) {
  using Ord = OrdKey<T>;
  constexpr int kSubs = kTile / SUBREC;
  const int lane = threadIdx.x;
  // Work items: lin0 = (query tile, way) in XCD-aware order of the (spatially sorted)
  // query tiles, so that each XCD's L2 holds the point tiles of one region; the `ways`
  // items of one query tile are adjacent, so they share an XCD (and the bound they
  // exchange stays in its L2).  One block per item, or (work != null) persistent
  // blocks: the items are cut into kXcds ranges, one per XCD, and each block takes the
  // next item of its own XCD's range (blocks are dealt to the XCDs round robin), then
  // of the following ranges once that one is used up; every block leaves once all
  // ranges are used up.  The counters start at zero (cleared before the launch).
  const int64_t nblk = static_cast<int64_t>(gridDim.x) * gridDim.y;
  const int64_t hw = static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x;
  int xr = static_cast<int>(hw % kXcds), tries = 0;
  auto next_item = [&]() __attribute__((always_inline)) -> int64_t {
    const int64_t qn = nitems / kXcds, rn = nitems % kXcds;
    while (tries < kXcds) {
      const int64_t lo = xr * qn + (xr < rn ? xr : rn);
      const int64_t hi = lo + qn + (xr < rn ? 1 : 0);
      unsigned got = 0;
      if (lane == 0) got = atomicAdd(work + xr, 1u);
      got = __builtin_amdgcn_readfirstlane(got);  // wave-uniform (scalar) from here on
      if (lo + static_cast<int64_t>(got) < hi) return lo + static_cast<int64_t>(got);
      xr = xr + 1 == kXcds ? 0 : xr + 1;
      ++tries;
    }
    return -1;
  };
  auto walk = [&](const int64_t lin0) __attribute__((always_inline)) {
  // item numbers, query tiles and ways fit in 32 bits (the host checks the counts)
  int64_t b, q0;
  const uint32_t item = static_cast<uint32_t>(lin0);
  const uint32_t way = ways == 1 ? 0u : item % static_cast<uint32_t>(ways);
  uint32_t lin = ways == 1 ? item : item / static_cast<uint32_t>(ways);
  int part = -1;  // a part of a split tile: its 64 / kSplitParts queries only
  if (torder) {  // the walk order of the query tiles
    const int32_t code = torder[lin];
    lin = static_cast<uint32_t>(code) >> 3;
    if (code & 1) part = (code >> 1) & 3;
  }
  if (qtiles) {
    const int64_t t = lin;
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (qtiles[mid] <= t) lo = mid; else hi = mid - 1;
    }
    b = lo;
    q0 = cl.q_begin(b) + (t - qtiles[b]) * 64;
  } else {
    const uint32_t per = static_cast<uint32_t>((cl.Q + 63) / 64);  // query tiles per cloud
    const uint32_t bb = cl.B == 1 ? 0u : lin / per;
    b = bb;
    q0 = cl.q_begin(b) + static_cast<int64_t>(lin - bb * per) * 64;
  }
  const int64_t qend = cl.q_end(b);
  const int64_t pos = q0 + lane;
  constexpr int kPartLanes = 64 / kSplitParts;
  const bool valid = pos < qend && (part < 0 || lane / kPartLanes == part);
  int64_t row = 0;
  // lanes past the cloud's queries keep a finite dummy query and a bound of -inf (thr
  // below): they need no box and insert nothing
  T qx = T(0), qy = qx, qz = qx;
  if (valid) {
    row = qperm[pos];
    qx = q[row * 3 + 0];
    qy = q[row * 3 + 1];
    qz = q[row * 3 + 2];
  }

  TopList<T, KMAX> L;
  L.init();

  // this cloud's tiles [tb, te) and super-tiles [sb, se)
  // tile and super-tile numbers fit in 32 bits (the host checks the tile count); the
  // walk state is kept narrow to leave scalar registers for the point records
  int tb, te, sb, se;
  if (cl.p_off) {
    tb = static_cast<int>(ptiles[b]);
    te = static_cast<int>(ptiles[b + 1]);
    sb = static_cast<int>(stiles[b]);
    se = static_cast<int>(stiles[b + 1]);
  } else {
    tb = static_cast<int>(b * ptiles[1]);
    te = tb + static_cast<int>(ptiles[1]);
    sb = static_cast<int>(b * stiles[1]);
    se = sb + static_cast<int>(stiles[1]);
  }
  // seed: last tile whose first code <= the code of the wave's middle query (scalar
  // search): the walk then starts nearer to most lanes' own neighbourhoods
  const uint32_t code0 =
      static_cast<uint32_t>(qkeys[min(q0 + (part < 0 ? 32 : kPartLanes * part + kPartLanes / 2), qend - 1)] &
                            0xffffffffu);
  int lo = tb, hi = te - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (__float_as_uint(hdr[static_cast<int64_t>(mid) * 8 + 7]) <= code0) lo = mid; else hi = mid - 1;
  }
  const int seed_t = lo;
  const int seed_s = sb + (seed_t - tb) / super;
  const int iway = static_cast<int>(way);

  // Split walk (ways > 1): this block owns the tiles t with (t - tb) % ways == way and
  // prunes, and filters insertions, with thr = the smallest last-kept key any way has
  // published for the query (starting from the query's seed bound).
  // Every way's list is a subset's top-KMAX, so its last key bounds the full top-KMAX
  // from above: a tile skipped against thr holds no point of the final result, whatever
  // the timing of the exchange.
  T thr = valid ? (ways > 1 ? Ord::dec(bnd[pos]) : static_cast<T>(__builtin_inf())) : static_cast<T>(-__builtin_inf());
  unsigned n_super = 0, n_tile = 0, n_scan = 0, n_pts = 0;
  Outward so(sb, se, seed_s);
  for (int sstep = 0; sstep < se - sb; ++sstep) {
    const int s = so.next(sstep);
    ++n_super;
    if (ways > 1 && valid)
      thr = fmin(thr, Ord::dec(__hip_atomic_load(bnd + pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
    if (!box_any<T>(load_hdr(shdr + static_cast<int64_t>(s) * 8), qx, qy, qz, fmin(L.last(), thr)))
      continue;
    const int t0 = tb + (s - sb) * super;
    const int t1 = min(te, t0 + super);
    Outward to(t0, t1, s == seed_s ? seed_t : t0);
    for (int tstep = 0; tstep < t1 - t0; ++tstep) {
      const int t = to.next(tstep);
      if ((t - tb) % ways != iway) continue;
      ++n_tile;
      if (!box_any<T>(load_hdr(hdr + static_cast<int64_t>(t) * 8), qx, qy, qz, fmin(L.last(), thr)))
        continue;
      ++n_scan;
      const int cnt = tcnt[t];
      // the tile's sub-tiles of SUBREC records: scan those some lane still needs
      const float* __restrict__ sh = sub + static_cast<int64_t>(t) * (kSubsMax * 8);
      // lim = fmin(last kept key, thr): thr is fixed within a tile, so lim changes only
      // with an insertion
      T lim = fmin(L.last(), thr);
      // headers four at a time (their scalar loads issued together, one wait); sub-tiles
      // past the tile's records (whole padding) are masked off
      constexpr int kBatch = kSubs < 4 ? kSubs : 4;
      unsigned m = 0;
#pragma unroll
      for (int u0 = 0; u0 < kSubs; u0 += kBatch) {
        BoxHdr hb[kBatch];
#pragma unroll
        for (int v = 0; v < kBatch; ++v) hb[v] = load_hdr(sh + (u0 + v) * 8);
#pragma unroll
        for (int v = 0; v < kBatch; ++v)
          if (box_any<T>(hb[v], qx, qy, qz, lim)) m |= 1u << (u0 + v);
      }
      const int live = (cnt + SUBREC - 1) / SUBREC;
      if (live < 32) m &= (1u << live) - 1u;
      if (m == 0) continue;
      const int total = SUBREC * __builtin_popcount(m);
      n_pts += total;
      const T* __restrict__ base = rec + static_cast<int64_t>(t) * (kTile * 16);
      // The needed sub-tiles' records a pair at a time, as one stream; the next pair's
      // scalar loads are issued before the current pair is evaluated.  Scalar loads
      // complete out of order, so each wait is for everything issued: with this order
      // it falls after the evaluation.  Sub-tiles are whole (padding records are NaN).
      // The pairs of a sub-tile are unrolled, so their offsets are constants and the
      // stream's bookkeeping runs once per sub-tile.
      unsigned rest = m;
      const T* cur = base + __builtin_ctz(rest) * (SUBREC * 16);
      rest &= rest - 1;
      Rec16<T> a0 = load_rec16(cur);
      Rec16<T> a1 = load_rec16(cur + 16);
      while (true) {
        // the next sub-tile (past the end: the current one again, loaded, not evaluated)
        const T* nxt = rest ? base + __builtin_ctz(rest) * (SUBREC * 16) : cur;
#pragma unroll
        for (int j = 0; j < SUBREC; j += 2) {
          const T* nx = j + 2 < SUBREC ? cur + (j + 2) * 16 : nxt;
          const Rec16<T> b0 = load_rec16(nx);
          const Rec16<T> b1 = load_rec16(nx + 16);
          const T c0 = rec_key<T, KEY>(qx, qy, qz, a0);
          const T c1 = rec_key<T, KEY>(qx, qy, qz, a1);
          // a key above the shared bound is above the final K-th key: never inserted
          if (c0 <= lim) {
            L.insert(c0, rec_index(a0.v[15]));
            lim = fmin(L.last(), thr);
          }
          if (c1 <= lim) {
            L.insert(c1, rec_index(a1.v[15]));
            lim = fmin(L.last(), thr);
          }
          a0 = b0;
          a1 = b1;
        }
        if (!rest) break;
        rest &= rest - 1;
        cur = nxt;
      }
      if (ways > 1 && valid && L.last() < thr) {  // publish an improved bound
        const T mine = L.last();
        const typename Ord::U old = atomicMin(bnd + pos, Ord::enc(mine));
        thr = fmin(mine, Ord::dec(old));
      }
    }
  }

  // The FP32 list holds a -0 key (possible only with sign-bit scales) as +0: a kept key
  // of zero is recomputed from the point arrays, which gives the scan's exact bits.
  const int64_t pb = cl.p_begin(b);
  auto out_key = [&](int s) -> T {
    T k = L.key(s);
    if constexpr (sizeof(T) == 4 && KEY == AKNN_KEY_ANISO) {
      const int n = L.id(s);
      if (k == T(0) && n != kNoPoint) {
        const int64_t r = pb + n;
        k = pair_key<T, KEY>(qx, qy, qz, pts[r * 3], pts[r * 3 + 1], pts[r * 3 + 2], rot[r * 9],
                             rot[r * 9 + 1], rot[r * 9 + 2], rot[r * 9 + 3], rot[r * 9 + 4],
                             rot[r * 9 + 5], rot[r * 9 + 6], rot[r * 9 + 7], rot[r * 9 + 8], sig[r * 3],
                             sig[r * 3 + 1], sig[r * 3 + 2]);
      }
    }
    return k;
  };
  if (valid) {
    if (ways == 1) {
      int last = 0;
#pragma unroll
      for (int s = 0; s < KMAX; ++s) {
        if (s < K) {
          dists[row * K + s] = out_key(s);
          idxs[row * K + s] = static_cast<int64_t>(L.id(s));
          last = L.id(s);
        }
      }
      if (last == kNoPoint) fill_unfilled_row<T>(dists + row * K, idxs + row * K, K);
    } else {
      T* lk = lkey + (pos * ways + way) * K;
      int32_t* li = lid + (pos * ways + way) * K;
#pragma unroll
      for (int s = 0; s < KMAX; ++s) {
        if (s < K) {
          lk[s] = out_key(s);
          li[s] = L.id(s);
        }
      }
    }
  }
  if (stats && lane == 0) {
    atomicAdd(stats + 0, static_cast<unsigned long long>(n_super));
    atomicAdd(stats + 1, static_cast<unsigned long long>(n_tile));
    atomicAdd(stats + 2, static_cast<unsigned long long>(n_scan));
    atomicAdd(stats + 3, static_cast<unsigned long long>(n_pts));
    atomicAdd(stats + 4, 1ull);
  }
  };
  for (int64_t item = work ? next_item() : xcd_block(hw, nblk); item >= 0; item = work ? next_item() : -1)
    walk(item);
}

// Super-tile headers: box union and min mu (negative if any tile is "always scan")
// of `super` consecutive tiles of one cloud.  One thread per super-tile.
__global__ void super_build_kernel(Clouds cl, const int64_t* __restrict__ ptiles,
                                   const int64_t* __restrict__ stiles, int64_t nsuper,
                                   const float* __restrict__ hdr, float* __restrict__ shdr, int super) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nsuper) return;
  int64_t b = 0, tb, te, sb;
  if (cl.p_off) {
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (stiles[mid] <= s) lo = mid; else hi = mid - 1;
    }
    b = lo;
    tb = ptiles[b];
    te = ptiles[b + 1];
    sb = stiles[b];
  } else {
    b = s / stiles[1];
    tb = b * ptiles[1];
    te = tb + ptiles[1];
    sb = b * stiles[1];
  }
  const int64_t t0 = tb + (s - sb) * super;
  const int64_t t1 = min(te, t0 + super);
  float lo3[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float hi3[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  float mu = __builtin_inff();
  bool always = false;
  for (int64_t t = t0; t < t1; ++t) {
    const float* h = hdr + t * 8;
    for (int c = 0; c < 3; ++c) {
      lo3[c] = fminf(lo3[c], h[c]);
      hi3[c] = fmaxf(hi3[c], h[3 + c]);
    }
    if (h[6] < 0.f) always = true;
    mu = fminf(mu, h[6]);
  }
  float* o = shdr + s * 8;
  for (int c = 0; c < 3; ++c) {
    o[c] = lo3[c];
    o[3 + c] = hi3[c];
  }
  o[6] = always ? -1.f : mu;
  o[7] = 0.f;
}

// Split walk: the shared bound of each sorted query position starts at the K-th
// smallest key over the 64 points of the tile its Morton code falls in (+inf if that
// tile holds fewer than K finite keys).  Those are real points of the query's cloud, so
// the value bounds the final K-th key from above and every tile the walk skips against
// it holds no point of the result; the walk itself inserts nothing from here (no
// duplicates).  Queries far from the walk's seed tile then prune from the first step
// instead of scanning with an infinite bound until the walk reaches them.
template <typename T, int KMAX, int KEY>
__global__ __launch_bounds__(256) void seed_bound_kernel(
    const T* __restrict__ q, Clouds cl, const uint32_t* __restrict__ qperm,
    const uint64_t* __restrict__ qkeys, const int64_t* __restrict__ ptiles, const T* __restrict__ rec,
    const float* __restrict__ hdr, int K, int64_t qrows, typename OrdKey<T>::U* __restrict__ bnd) {
  using Ord = OrdKey<T>;
  const int64_t pos = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pos >= qrows) return;
  const int64_t b = cl.cloud_of_query(pos);
  int64_t tb, te;
  if (cl.p_off) {
    tb = ptiles[b];
    te = ptiles[b + 1];
  } else {
    tb = b * ptiles[1];
    te = tb + ptiles[1];
  }
  T bound = __builtin_inf();
  if (te > tb) {
    const uint32_t code = static_cast<uint32_t>(qkeys[pos] & 0xffffffffu);
    int64_t lo = tb, hi = te - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (__float_as_uint(hdr[mid * 8 + 7]) <= code) lo = mid; else hi = mid - 1;
    }
    const int64_t row = qperm[pos];
    const T qx = q[row * 3 + 0], qy = q[row * 3 + 1], qz = q[row * 3 + 2];
    T key[KMAX];
#pragma unroll
    for (int s = 0; s < KMAX; ++s) key[s] = __builtin_inf();
    const T* __restrict__ base = rec + lo * (kTile * 16);
#pragma unroll 4
    for (int j = 0; j < kTile; ++j) {
      const T c = rec_key<T, KEY>(qx, qy, qz, load_rec16(base + j * 16));
      if (c < key[KMAX - 1]) {  // sorted insert of the value (NaN never enters)
        T carry = c;
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
          const T lo_v = fmin(key[s], carry);
          carry = fmax(key[s], carry);
          key[s] = lo_v;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < KMAX; ++s)
      if (s == K - 1) bound = key[s];
  }
  bnd[pos] = Ord::enc(bound);
}

// Split walk: merge the `ways` per-way top-K lists of each sorted query position.
// A wave holds 64 / ways queries, one lane per (query, way); a butterfly over the
// ways merges the lists (disjoint tile sets, so no duplicates), lane way 0 writes.
template <typename T, int KMAX>
__global__ __launch_bounds__(64) void pruned_merge_kernel(const T* __restrict__ lkey,
                                                          const int32_t* __restrict__ lid,
                                                          const uint32_t* __restrict__ qperm,
                                                          int64_t qrows, int ways, int K,
                                                          T* __restrict__ dists,
                                                          int64_t* __restrict__ idxs) {
  const int lane = threadIdx.x;
  const int per = 64 / ways;
  const int w = lane % ways;
  const int64_t pos = static_cast<int64_t>(blockIdx.x) * per + lane / ways;
  const bool ok = pos < qrows;
  T key[KMAX];
  int id[KMAX];
#pragma unroll
  for (int s = 0; s < KMAX; ++s) {
    key[s] = __builtin_inf();
    id[s] = kTileSentinel;
  }
  if (ok) {
    const T* lk = lkey + (pos * ways + w) * K;
    const int32_t* li = lid + (pos * ways + w) * K;
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      if (s < K) {
        key[s] = lk[s];
        id[s] = li[s];
      }
    }
  }
  for (int off = 1; off < ways; off <<= 1) {
    T pk[KMAX];
    int pi[KMAX];
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      pk[s] = __shfl_xor(key[s], off, 64);
      pi[s] = __shfl_xor(id[s], off, 64);
    }
#pragma unroll
    for (int s = 0; s < KMAX; ++s)
      if ((pk[s] < key[KMAX - 1]) | ((pk[s] == key[KMAX - 1]) & (pi[s] < id[KMAX - 1])))
        lex_insert<KMAX>(key, id, pk[s], pi[s]);
  }
  if (ok && w == 0) {
    const int64_t row = qperm[pos];
    int last = 0;
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      if (s < K) {
        dists[row * K + s] = key[s];
        idxs[row * K + s] = static_cast<int64_t>(id[s]);
        last = id[s];
      }
    }
    if (last == kNoPoint) fill_unfilled_row<T>(dists + row * K, idxs + row * K, K);
  }
}

// ------------------------------------------------------------------ per-query walk
// Sparse query sets (few queries per cloud point): the 64 queries of a Morton tile lie
// far apart, so the tile walk scans the union of their neighbourhoods.  Here one wave
// serves one query: it evaluates a whole point tile per step (lane j: record j), keeps
// the query's top-K as a sorted (key, index) list in lanes 0..K-1, and visits
// super-tiles and tiles best first (smallest certified lower bound first; "always
// scan" boxes before all), each at most once, skipping those whose bound exceeds the
// list's last key.  The result is the same exact top-K, in the same order.

template <typename T>
__device__ __forceinline__ T lane_value(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long x = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(x), l));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(x >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
  }
}

// DPP move: lanes whose source is outside the pattern, or that the row / bank masks
// leave out, keep `old`.
template <int CTRL, int ROW_MASK, int BANK_MASK, typename T>
__device__ __forceinline__ T dpp_move(T old, T v) {
  if constexpr (sizeof(T) == 4) {
    int o, x;
    __builtin_memcpy(&o, &old, 4);
    __builtin_memcpy(&x, &v, 4);
    const int r = __builtin_amdgcn_update_dpp(o, x, CTRL, ROW_MASK, BANK_MASK, false);
    T out;
    __builtin_memcpy(&out, &r, 4);
    return out;
  } else {
    int o[2], x[2], r[2];
    __builtin_memcpy(o, &old, 8);
    __builtin_memcpy(x, &v, 8);
    r[0] = __builtin_amdgcn_update_dpp(o[0], x[0], CTRL, ROW_MASK, BANK_MASK, false);
    r[1] = __builtin_amdgcn_update_dpp(o[1], x[1], CTRL, ROW_MASK, BANK_MASK, false);
    T out;
    __builtin_memcpy(&out, r, 8);
    return out;
  }
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr3 = 0x113, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

// Minimum over the wave of NaN-free values (`inf`: +infinity held in a register), with
// DPP row shifts and broadcasts: lane 63 ends with the minimum.
template <typename T>
__device__ __forceinline__ T wave_min(T x, T inf) {
  auto mn = [](T a, T b) { return b < a ? b : a; };
  T a = mn(x, dpp_move<kDppRowShr1, 0xf, 0xf>(inf, x));
  a = mn(a, dpp_move<kDppRowShr2, 0xf, 0xf>(inf, x));
  a = mn(a, dpp_move<kDppRowShr3, 0xf, 0xf>(inf, x));
  a = mn(a, dpp_move<kDppRowShr4, 0xf, 0xe>(inf, a));
  a = mn(a, dpp_move<kDppRowShr8, 0xf, 0xc>(inf, a));
  a = mn(a, dpp_move<kDppRowBcast15, 0xa, 0xf>(inf, a));
  a = mn(a, dpp_move<kDppRowBcast31, 0xc, 0xf>(inf, a));
  return lane_value<T>(a, 63);
}

// Lowest lane among `sel` holding the smallest `v` (NaN ordered first); some lane must
// be selected.
template <typename T>
__device__ __forceinline__ int wave_argmin(T v, bool sel, T inf) {
  const T x = sel ? (v != v ? -inf : v) : inf;
  const T m = wave_min<T>(x, inf);
  unsigned long long hit = __ballot(sel && x == m);
  if (hit == 0) hit = __ballot(sel);  // never taken when the minimum is exact; stays in range
  return __builtin_ctzll(hit);
}

// Certified lower bound of a box's keys for the query (negative infinity: "always
// scan" box, whose mu is negative).
template <typename T>
__device__ __forceinline__ T box_bound(const float* __restrict__ h, T qx, T qy, T qz) {
  const float4 a = *reinterpret_cast<const float4*>(h);
  const float4 c = *reinterpret_cast<const float4*>(h + 4);
  return c.z < 0.f ? -__builtin_inf() : box_lower_bound<T>(a.x, a.y, a.z, a.w, c.x, c.y, c.z, qx, qy, qz);
}

template <typename T>
__device__ __forceinline__ bool lex_less(T a, int ia, T b, int ib) {
  return (a < b) | ((a == b) & (ia < ib));
}

// Value of lane ^ J: DPP quad permutes (J = 1, 2) and row shifts (J = 4, 8), a
// ds_swizzle (J = 16), a permute through LDS (J = 32).
constexpr int kDppQuadSwap1 = 0xB1, kDppQuadSwap2 = 0x4E, kDppRowShl4 = 0x104, kDppRowShl8 = 0x108;
constexpr int kSwizzleXor16 = 0x401F;  // bitmask mode: and 0x1f, xor 0x10

template <int J, typename T>
__device__ __forceinline__ T lane_xor(T v, int lane) {
  if constexpr (J == 1) {
    return dpp_move<kDppQuadSwap1, 0xf, 0xf>(v, v);
  } else if constexpr (J == 2) {
    return dpp_move<kDppQuadSwap2, 0xf, 0xf>(v, v);
  } else if constexpr (J == 4) {
    const T up = dpp_move<kDppRowShl4, 0xf, 0xf>(v, v), dn = dpp_move<kDppRowShr4, 0xf, 0xf>(v, v);
    return (lane & 4) ? dn : up;
  } else if constexpr (J == 8) {
    const T up = dpp_move<kDppRowShl8, 0xf, 0xf>(v, v), dn = dpp_move<kDppRowShr8, 0xf, 0xf>(v, v);
    return (lane & 8) ? dn : up;
  } else if constexpr (J == 16) {
    if constexpr (sizeof(T) == 4) {
      int x;
      __builtin_memcpy(&x, &v, 4);
      x = __builtin_amdgcn_ds_swizzle(x, kSwizzleXor16);
      T out;
      __builtin_memcpy(&out, &x, 4);
      return out;
    } else {
      int x[2];
      __builtin_memcpy(x, &v, 8);
      x[0] = __builtin_amdgcn_ds_swizzle(x[0], kSwizzleXor16);
      x[1] = __builtin_amdgcn_ds_swizzle(x[1], kSwizzleXor16);
      T out;
      __builtin_memcpy(&out, x, 8);
      return out;
    }
  } else {
    static_assert(J == 32, "lane_xor: J in 1..32");
    return __shfl_xor(v, 32, 64);
  }
}

// Compare-exchange with lane ^ J in the (key, index) order: the lane keeps the smaller
// of the pair if keep_min, else the larger.
template <int J, typename T>
__device__ __forceinline__ void cmp_exchange(T& k, int& i, int lane, bool keep_min) {
  const T ok = lane_xor<J>(k, lane);
  const int oi = lane_xor<J>(i, lane);
  const bool take = keep_min ? lex_less(ok, oi, k, i) : lex_less(k, i, ok, oi);
  k = take ? ok : k;
  i = take ? oi : i;
}

// Bitonic sort of the wave's 64 (key, index) pairs, ascending by lane: stage KK, steps
// J = KK/2 .. 1.
template <int KK, int J, typename T>
__device__ __forceinline__ void bitonic_steps(T& k, int& i, int lane) {
  cmp_exchange<J>(k, i, lane, ((lane & J) == 0) == ((lane & KK) == 0));
  if constexpr (J > 1) bitonic_steps<KK, J / 2>(k, i, lane);
}
template <int KK, typename T>
__device__ __forceinline__ void bitonic_sort64(T& k, int& i, int lane) {
  bitonic_steps<KK, KK / 2>(k, i, lane);
  if constexpr (KK < 64) bitonic_sort64<KK * 2>(k, i, lane);
}

// Ascending half-cleaner steps J .. 1 (orders a bitonic run of 2J lanes).
template <int J, typename T>
__device__ __forceinline__ void bitonic_merge(T& k, int& i, int lane) {
  cmp_exchange<J>(k, i, lane, (lane & J) == 0);
  if constexpr (J > 1) bitonic_merge<J / 2>(k, i, lane);
}

template <typename T, int KEY>
__global__ __launch_bounds__(256) void knn_wave_kernel(
    const T* __restrict__ q, Clouds cl, int64_t qrows, const int64_t* __restrict__ ptiles,
    const int64_t* __restrict__ stiles, const T* __restrict__ rec, const float* __restrict__ hdr,
    const float* __restrict__ shdr, int K, T* __restrict__ dists, int64_t* __restrict__ idxs,
    unsigned long long* __restrict__ stats, int super) {
  const int lane = threadIdx.x & 63;
  // the wave's query (made wave-uniform, so the walk's bookkeeping is scalar)
  const int64_t row = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) +
                      __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  if (row >= qrows) return;  // whole waves
  const int64_t b = cl.cloud_of_query(row);
  int64_t tb, te, sb, se;
  if (cl.p_off) {
    tb = ptiles[b];
    te = ptiles[b + 1];
    sb = stiles[b];
    se = stiles[b + 1];
  } else {
    tb = b * ptiles[1];
    te = tb + ptiles[1];
    sb = b * stiles[1];
    se = sb + stiles[1];
  }
  const T qx = q[row * 3 + 0], qy = q[row * 3 + 1], qz = q[row * 3 + 2];
  // +inf reaches the wave-uniform values through a vector register: for a double kept
  // in scalar registers hipcc (gfx950) emits s_mov_b64 with a 64-bit literal of which
  // the encoding keeps the low half (+inf became 0.0)
  T inf_v = __builtin_inf();
  asm volatile("" : "+v"(inf_v));
  // lanes 0..Kp-1 (K rounded up to a power of two): the sorted list; lane K-1 holds
  // the pruning bound, lanes K..Kp-1 pad the bitonic merge
  int Kp = 1;
  while (Kp < K) Kp <<= 1;
  T lk = inf_v;
  int li = kTileSentinel;
  T last_k = lane_value<T>(inf_v, 0);
  int last_i = kTileSentinel;
  unsigned n_super = 0, n_tile = 0, n_scan = 0;

  const int64_t nsup = se - sb;
  const int64_t nchunk = (nsup + 63) / 64;
  // chunks of 64 super-tiles: the one holding the closest super-tile first, then
  // outward (a cloud of up to 64K points is one chunk)
  int64_t first = 0;
  if (nchunk > 1) {
    T best = lane_value<T>(inf_v, 0);
    int64_t bc = 0;
    for (int64_t c = 0; c < nchunk; ++c) {
      const int64_t s = sb + c * 64 + lane;
      const T v = s < se ? box_bound<T>(shdr + s * 8, qx, qy, qz) : __builtin_inf();
      const int l = wave_argmin<T>(v, s < se, inf_v);
      const T m = lane_value<T>(v, l);
      if (m < best) {
        best = m;
        bc = c;
      }
    }
    first = bc;
  }
  // One tile's 64 records against the list (lane j: record j).
  auto consume = [&](const Rec16<T>& r) {
    const T ck = rec_key<T, KEY>(qx, qy, qz, r);
    const int ci = rec_index(r.v[15]);
    const bool cand = lex_less(ck, ci, last_k, last_i);
    unsigned long long m = __ballot(cand);
    if (m) ++n_scan;
    if (__builtin_popcountll(m) > kWaveSerialInserts) {
      // many candidates (the first tiles): sort the tile's 64 (key, index) pairs
      // across the wave, then merge the K smallest into the list (the elementwise
      // minimum of the list and the reversed candidates is bitonic and holds the Kp
      // smallest of both; a bitonic merge orders it)
      T sk = cand ? ck : inf_v;
      int si = cand ? ci : kTileSentinel;
      bitonic_sort64<2>(sk, si, lane);
      const T rk = __shfl(sk, (Kp - 1 - lane) & 63, 64);
      const int ri = __shfl(si, (Kp - 1 - lane) & 63, 64);
      if (lane < Kp && lex_less(rk, ri, lk, li)) {
        lk = rk;
        li = ri;
      }
      switch (Kp) {
        case 64: bitonic_merge<32>(lk, li, lane); break;
        case 32: bitonic_merge<16>(lk, li, lane); break;
        case 16: bitonic_merge<8>(lk, li, lane); break;
        case 8: bitonic_merge<4>(lk, li, lane); break;
        case 4: bitonic_merge<2>(lk, li, lane); break;
        case 2: bitonic_merge<1>(lk, li, lane); break;
        default: break;
      }
      last_k = lane_value<T>(lk, K - 1);
      last_i = __builtin_amdgcn_readlane(li, K - 1);
      m = 0;
    }
    while (m) {
      const int l = __builtin_ctzll(m);
      m &= m - 1;
      const T k = lane_value<T>(ck, l);
      const int i = __builtin_amdgcn_readlane(ci, l);
      if (!lex_less(k, i, last_k, last_i)) continue;
      // insert into the sorted list: lanes at or after the insertion point shift up
      const bool lt = (lane < Kp) & lex_less(k, i, lk, li);
      const T pk = dpp_move<kDppWaveShr1, 0xf, 0xf>(lk, lk);
      const int pi = dpp_move<kDppWaveShr1, 0xf, 0xf>(li, li);
      const bool plt = dpp_move<kDppWaveShr1, 0xf, 0xf>(0, static_cast<int>(lt)) != 0;
      if (lt) {
        lk = plt ? pk : k;
        li = plt ? pi : i;
      }
      last_k = lane_value<T>(lk, K - 1);
      last_i = __builtin_amdgcn_readlane(li, K - 1);
    }
  };

  int64_t up = first, down = first - 1;
  for (int64_t step = 0; step < nchunk; ++step) {
    const bool take_up = (up < nchunk) && ((step & 1) == 0 || down < 0);
    const int64_t c = take_up ? up++ : down--;
    const int64_t s = sb + c * 64 + lane;
    const bool sin = s < se;
    const T slb = sin ? box_bound<T>(shdr + s * 8, qx, qy, qz) : inf_v;
    bool sdone = !sin;
    while (true) {
      const bool ssel = !sdone && !(slb > last_k);
      if (!__any(ssel)) break;
      const int sl = wave_argmin<T>(slb, ssel, inf_v);
      if (lane == sl) sdone = true;
      ++n_super;
      const int64_t t0 = tb + (c * 64 + sl) * super;
      const int64_t nt = te - t0 < super ? te - t0 : super;
      const bool tin = lane < nt;
      const T tlb = tin ? box_bound<T>(hdr + (t0 + lane) * 8, qx, qy, qz) : inf_v;
      bool tdone = !tin;
      while (true) {
        const bool tsel = !tdone && !(tlb > last_k);
        if (!__any(tsel)) break;
        const int tl = wave_argmin<T>(tlb, tsel, inf_v);
        if (lane == tl) tdone = true;
        ++n_tile;
        consume(load_rec16(rec + (t0 + tl) * (kTile * 16) + lane * 16));
      }
    }
  }
  if (last_i == kNoPoint) {
    // unfilled tail (fewer than K keys compare): the lowest point indices not in the
    // list, NaN keys, as fill_unfilled_row does for the other searches
    bool used = false;
    for (int j = 0; j < K; ++j) used |= __builtin_amdgcn_readlane(li, j) == lane;
    unsigned long long free_idx = __ballot(!used && lane < K);
    const int filled = __builtin_popcountll(__ballot(lane < K && li != kNoPoint));
    if (lane < K && li == kNoPoint) {
      for (int t = filled; t < lane; ++t) free_idx &= free_idx - 1;
      li = __builtin_ctzll(free_idx);
      lk = static_cast<T>(__builtin_nan(""));
    }
  }
  if (lane < K) {
    dists[row * K + lane] = lk;
    idxs[row * K + lane] = static_cast<int64_t>(li);
  }
  if (stats && lane == 0) {
    atomicAdd(stats + 0, static_cast<unsigned long long>(n_super));
    atomicAdd(stats + 1, static_cast<unsigned long long>(n_tile));
    atomicAdd(stats + 2, static_cast<unsigned long long>(n_scan));
    atomicAdd(stats + 3, static_cast<unsigned long long>(n_tile) * kTile);
    atomicAdd(stats + 4, 1ull);
  }
}

template <typename T, int KEY>
static hipError_t launch_wave(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  if (L.qrows == 0) return hipSuccess;
  hipLaunchKernelGGL((knn_wave_kernel<T, KEY>), dim3(static_cast<unsigned>((L.qrows + 3) / 4)), dim3(256), 0,
                     st, static_cast<const T*>(L.q), L.cl, L.qrows, S.ptp, S.stp, static_cast<const T*>(S.rec),
                     S.hdr, S.shdr, static_cast<int>(L.K), static_cast<T*>(L.dists), L.idxs, L.stats,
                     static_cast<int>(L.super_tiles));
  return hipGetLastError();
}

// Blocks of `threads` threads of kernel `fn` resident on the current device at once (0 if
// unknown), cached per kernel and device.
static int64_t resident_blocks(const void* fn, int threads) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int64_t> cache;
  std::lock_guard<std::mutex> guard(mu);
  const auto it = cache.find({fn, dev});
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const int64_t r = static_cast<int64_t>(per_cu) * cus;
  cache[{fn, dev}] = r;
  return r;
}

__global__ void zero_u32_kernel(unsigned* __restrict__ p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0u;
}

template <typename T, int KMAX, int KEY, int SUBREC, int WAVES>
static void launch_search(const PrunedLaunch& L, const PrunedState& S, dim3 grid, int ways, hipStream_t st) {
  using U = typename OrdKey<T>::U;
  auto kern = knn_pruned_kernel<T, KMAX, KEY, SUBREC, WAVES>;
  const int64_t nitems = static_cast<int64_t>(grid.x) * grid.y;
  unsigned* work = nullptr;
  if (L.persistent) {  // as many blocks as fit on the device at once, over the work items
    const int64_t res = resident_blocks(reinterpret_cast<const void*>(kern), 64);
    if (res > 0 && res < nitems) {
      grid = dim3(static_cast<unsigned>(res));
      work = S.work;
      hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, work, kXcds);
    }
  }
  hipLaunchKernelGGL(kern, grid, dim3(64), 0, st,
                     static_cast<const T*>(L.q), L.cl, L.cl.q_off ? S.qtp : nullptr, S.qperm, S.qkeys_s, S.torder,
                     S.ptp,
                     S.stp, static_cast<const T*>(S.rec), S.hdr, S.shdr, S.sub, S.tcnt, static_cast<int>(L.K),
                     static_cast<T*>(L.dists), L.idxs, L.stats, ways, static_cast<T*>(S.lkey), S.lid,
                     static_cast<U*>(S.bnd), static_cast<const T*>(L.p), static_cast<const T*>(L.sigma),
                     static_cast<const T*>(L.R), static_cast<int>(L.super_tiles), work, nitems);
}

template <typename T, int KMAX, int KEY>
static hipError_t launch_pruned_k(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  using U = typename OrdKey<T>::U;
  const int ways = L.ways > 1 ? L.ways : 1;
  if (ways > 1)
    hipLaunchKernelGGL((seed_bound_kernel<T, KMAX, KEY>), dim3(static_cast<unsigned>((L.qrows + 255) / 256)),
                       dim3(256), 0, st, static_cast<const T*>(L.q), L.cl, S.qperm, S.qkeys_s, S.ptp,
                       static_cast<const T*>(S.rec), S.hdr, static_cast<int>(L.K), L.qrows,
                       static_cast<U*>(S.bnd));
  // walk items (query tiles, and the parts of split ones) x ways, one dimension
  const dim3 grid(static_cast<unsigned>(walk_items(L.nqtiles) * ways), 1);
  // tile geometry and occupancy variants (the host admits them for FP32, 9..16 neighbours)
  if constexpr (sizeof(T) == 4 && KMAX == 16) {
    if (L.sub_rec == 4) launch_search<T, KMAX, KEY, 4, 0>(L, S, grid, ways, st);
    else if (L.sub_rec == 16) launch_search<T, KMAX, KEY, 16, 0>(L, S, grid, ways, st);
    else if (L.waves == 4) launch_search<T, KMAX, KEY, 8, 4>(L, S, grid, ways, st);
    else if (L.waves == 8) launch_search<T, KMAX, KEY, 8, 8>(L, S, grid, ways, st);
    else launch_search<T, KMAX, KEY, 8, 0>(L, S, grid, ways, st);
  } else {
    if (L.sub_rec != kSubRecDefault || L.waves != 0) return hipErrorInvalidValue;
    launch_search<T, KMAX, KEY, 8, 0>(L, S, grid, ways, st);
  }
  if (ways > 1) {
    const int per = 64 / ways;
    hipLaunchKernelGGL((pruned_merge_kernel<T, KMAX>), dim3(static_cast<unsigned>((L.qrows + per - 1) / per)),
                       dim3(64), 0, st, static_cast<const T*>(S.lkey), S.lid, S.qperm, L.qrows, ways,
                       static_cast<int>(L.K), static_cast<T*>(L.dists), L.idxs);
  }
  return hipGetLastError();
}

template <typename T, int KEY>
static hipError_t launch_pruned_topk(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  if (L.K <= 1) return launch_pruned_k<T, 1, KEY>(L, S, st);
  if (L.K <= 4) return launch_pruned_k<T, 4, KEY>(L, S, st);
  if (L.K <= 8) return launch_pruned_k<T, 8, KEY>(L, S, st);
  if (L.K <= 16) return launch_pruned_k<T, 16, KEY>(L, S, st);
  if (L.K <= 32) return launch_pruned_k<T, 32, KEY>(L, S, st);
  return launch_pruned_k<T, 64, KEY>(L, S, st);
}

static int key_bits64(int64_t B) {
  int bits = kCodeBits;  // Morton code below the cloud id
  while ((int64_t(1) << (bits - kCodeBits)) < B) ++bits;
  return bits;
}

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

// Workspace carve (each piece 256-B aligned).
struct Carver {
  char* base;
  size_t used = 0;
  char* take(size_t n) {
    char* p = base ? base + used : nullptr;
    used += al(n);
    return p;
  }
};

// The point index: the cloud-only pieces (boxes, point order, tile records and
// headers).  Carved from L.index when the caller keeps a prebuilt one (resident
// clouds), else from the workspace and rebuilt by every call.
static void carve_index(const PrunedLaunch& L, Carver& c, PrunedState& S) {
  const size_t el = L.dtype == AKNN_F64 ? 8 : 4;
  S.box = reinterpret_cast<float*>(c.take(L.cl.B * 8 * 4));
  S.obox = reinterpret_cast<unsigned*>(c.take(L.cl.B * 6 * 4));
  S.pperm = reinterpret_cast<uint32_t*>(c.take(L.prows * 4));
  S.ptp = reinterpret_cast<int64_t*>(c.take((L.cl.B + 1) * 8));
  S.stp = reinterpret_cast<int64_t*>(c.take((L.cl.B + 1) * 8));
  S.rec = c.take(L.ntiles * kTile * 16 * el);
  S.hdr = reinterpret_cast<float*>(c.take(L.ntiles * 8 * 4));
  S.shdr = reinterpret_cast<float*>(c.take(L.nsuper_alloc * 8 * 4));
  S.sub = reinterpret_cast<float*>(c.take(L.ntiles * kSubsMax * 8 * 4));
  S.tcnt = reinterpret_cast<int32_t*>(c.take(L.ntiles * 4));
}

size_t pruned_index_bytes(const PrunedLaunch& L) {
  Carver c{nullptr};
  PrunedState S{};
  carve_index(L, c, S);
  return c.used;
}

static PrunedState carve(const PrunedLaunch& L, size_t tmp_bytes) {
  PrunedState S{};
  Carver c{static_cast<char*>(L.ws)};
  const size_t el = L.dtype == AKNN_F64 ? 8 : 4;
  if (L.index) {
    Carver ci{static_cast<char*>(const_cast<void*>(L.index))};
    carve_index(L, ci, S);
  } else {
    carve_index(L, c, S);
  }
  S.pkeys = reinterpret_cast<uint64_t*>(c.take(L.prows * 8));
  S.pkeys_s = reinterpret_cast<uint64_t*>(c.take(L.prows * 8));
  S.qkeys = reinterpret_cast<uint64_t*>(c.take(L.qrows * 8));
  S.qkeys_s = reinterpret_cast<uint64_t*>(c.take(L.qrows * 8));
  S.qperm = reinterpret_cast<uint32_t*>(c.take(L.qrows * 4));
  S.qtp = reinterpret_cast<int64_t*>(c.take((L.cl.B + 1) * 8));
  S.work = reinterpret_cast<unsigned*>(c.take(kXcds * 4));
  S.torder = reinterpret_cast<int32_t*>(c.take(walk_items(L.nqtiles) * 4));
  S.tmp = c.take(tmp_bytes);
  S.tmp_bytes = tmp_bytes;
  if (L.ways > 1) {
    const size_t lists = static_cast<size_t>(L.qrows) * L.ways * L.K;
    S.lkey = c.take(lists * el);
    S.lid = reinterpret_cast<int32_t*>(c.take(lists * 4));
    S.bnd = c.take(static_cast<size_t>(L.qrows) * el);
  }
  S.total = c.used;
  return S;
}

static hipError_t sort_temp(const PrunedLaunch& L, size_t* bytes) {
  const int64_t n = L.prows > L.qrows ? L.prows : L.qrows;
  *bytes = radix_sort_workspace(n, 8);
  return hipSuccess;
}

hipError_t pruned_workspace(const PrunedLaunch& L, size_t* bytes) {
  size_t tmp = 0;
  const hipError_t e = sort_temp(L, &tmp);
  if (e != hipSuccess) return e;
  PrunedLaunch probe = L;
  probe.ws = nullptr;
  probe.ways = L.ways_max > L.ways ? L.ways_max : L.ways;
  *bytes = carve(probe, tmp).total;
  return hipSuccess;
}

__global__ void tile_prefix_kernel2(const int64_t* __restrict__ off, int64_t B, int64_t per,
                                    int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  for (int64_t b = 0; b < B; ++b) {
    out[b] = acc;
    acc += (off[b + 1] - off[b] + per - 1) / per;
  }
  out[B] = acc;
}

// super-tile prefix from the tile prefix (ragged), `super` tiles per super-tile
__global__ void super_prefix_kernel(const int64_t* __restrict__ ptp, int64_t B, int64_t super,
                                    int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  for (int64_t b = 0; b < B; ++b) {
    out[b] = acc;
    acc += (ptp[b + 1] - ptp[b] + super - 1) / super;
  }
  out[B] = acc;
}

__global__ void dense_prefix_kernel(int64_t per_cloud, int64_t* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = 0;
    out[1] = per_cloud;
  }
}

// The point index: boxes, point Morton codes and order, tile records/headers and
// super-tile headers.
template <typename T, int KEY>
static hipError_t prep_points(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  const int bits = key_bits64(L.cl.B);
  const T* p = static_cast<const T*>(L.p);
  const int bs = 256;
  hipLaunchKernelGGL(box_init_kernel, dim3(static_cast<unsigned>((L.cl.B * 6 + bs - 1) / bs)), dim3(bs), 0, st,
                     S.obox, L.cl.B);
  const int64_t chunks = (L.max_points + kBoxChunk - 1) / kBoxChunk;
  hipLaunchKernelGGL((cloud_box_kernel<T>), dim3(static_cast<unsigned>(chunks > 0 ? chunks : 1),
                     static_cast<unsigned>(L.cl.B)), dim3(256), 0, st, p, L.cl, S.obox);
  hipLaunchKernelGGL(box_final_kernel, dim3(static_cast<unsigned>((L.cl.B * 8 + bs - 1) / bs)), dim3(bs), 0, st,
                     S.obox, L.cl.B, S.box);
  hipLaunchKernelGGL((codes_kernel<T, false>), dim3(static_cast<unsigned>((L.prows + bs - 1) / bs)), dim3(bs), 0,
                     st, p, L.cl, L.prows, S.box, S.pkeys);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = radix_sort_pairs<uint64_t>(S.pkeys, S.pkeys_s, nullptr, S.pperm, L.prows, 0, bits, S.tmp, st);
  if (e != hipSuccess) return e;
  if (L.cl.p_off) {
    hipLaunchKernelGGL(tile_prefix_kernel2, dim3(1), dim3(64), 0, st, L.cl.p_off, L.cl.B, int64_t(kTile), S.ptp);
    hipLaunchKernelGGL(super_prefix_kernel, dim3(1), dim3(64), 0, st, S.ptp, L.cl.B,
                       static_cast<int64_t>(L.super_tiles), S.stp);
  } else {
    const int64_t per = (L.cl.P + kTile - 1) / kTile;
    hipLaunchKernelGGL(dense_prefix_kernel, dim3(1), dim3(64), 0, st, per, S.ptp);
    hipLaunchKernelGGL(dense_prefix_kernel, dim3(1), dim3(64), 0, st, (per + L.super_tiles - 1) / L.super_tiles,
                       S.stp);
  }
  const dim3 tg(static_cast<unsigned>(L.ntiles));
  const T* sg = static_cast<const T*>(L.sigma);
  const T* rt = static_cast<const T*>(L.R);
  if (L.sub_rec == 4)
    hipLaunchKernelGGL((tile_build_kernel<T, KEY, 4>), tg, dim3(64), 0, st, p, sg, rt, L.cl, S.ptp, S.pperm,
                       S.pkeys_s, static_cast<T*>(S.rec), S.hdr, S.sub, S.tcnt, L.mu_classes);
  else if (L.sub_rec == 16)
    hipLaunchKernelGGL((tile_build_kernel<T, KEY, 16>), tg, dim3(64), 0, st, p, sg, rt, L.cl, S.ptp, S.pperm,
                       S.pkeys_s, static_cast<T*>(S.rec), S.hdr, S.sub, S.tcnt, L.mu_classes);
  else
    hipLaunchKernelGGL((tile_build_kernel<T, KEY, 8>), tg, dim3(64), 0, st, p, sg, rt, L.cl, S.ptp, S.pperm,
                       S.pkeys_s, static_cast<T*>(S.rec), S.hdr, S.sub, S.tcnt, L.mu_classes);
  hipLaunchKernelGGL(super_build_kernel, dim3(static_cast<unsigned>((L.nsuper + bs - 1) / bs)), dim3(bs), 0, st,
                     L.cl, S.ptp, S.stp, L.nsuper, S.hdr, S.shdr, static_cast<int>(L.super_tiles));
  return hipGetLastError();
}

// Walk order of the query tiles.  A tile's walk scans the union of its 64 queries'
// neighbourhoods, so a tile whose queries spread over a wide region (its first and last
// Morton codes differ in a high bit: the tile straddles far-apart regions of the sorted
// order) takes several times the typical walk, and when such tiles start late they set
// the kernel's end.  Within each XCD's contiguous range of tiles (the ranges of
// xcd_block and of the persistent work lists), tiles are ordered by that bit, highest
// first (longest walks first); equal bits in any order (the results do not depend on
// the order).  One block per range.

__device__ __forceinline__ int tile_spread(const uint64_t* __restrict__ keys, const Clouds& cl,
                                           const int64_t* __restrict__ qtiles, int64_t t) {
  int64_t q0, qe;
  if (qtiles) {
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (qtiles[mid] <= t) lo = mid; else hi = mid - 1;
    }
    q0 = cl.q_begin(lo) + (t - qtiles[lo]) * 64;
    qe = cl.q_end(lo);
  } else {
    const int64_t per = (cl.Q + 63) / 64;
    const int64_t b = t / per;
    q0 = b * cl.Q + (t - b * per) * 64;
    qe = (b + 1) * cl.Q;
  }
  const int64_t q1 = q0 + 63 < qe ? q0 + 63 : qe - 1;
  const uint64_t x = keys[q0] ^ keys[q1];
  return x ? 64 - __builtin_clzll(x) : 0;  // 0..64
}

__global__ __launch_bounds__(1024) void tile_order_kernel(const uint64_t* __restrict__ keys, Clouds cl,
                                                          const int64_t* __restrict__ qtiles, int64_t ntiles,
                                                          int32_t* __restrict__ order) {
  __shared__ unsigned cnt[65];
  const int64_t qn = ntiles / kXcds, rn = ntiles % kXcds;
  const int64_t x = blockIdx.x;
  const int64_t lo = x * qn + (x < rn ? x : rn);
  const int64_t hi = lo + qn + (x < rn ? 1 : 0);
  const int64_t extra = split_extra(ntiles);  // this range's extra items
  const int64_t nsplit = extra / (kSplitParts - 1);
  const int64_t ilo = lo + x * extra;         // this range's first walk item
  for (int i = threadIdx.x; i < 65; i += blockDim.x) cnt[i] = 0u;
  __syncthreads();
  for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) atomicAdd(&cnt[tile_spread(keys, cl, qtiles, t)], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {  // bin starts, highest spread first
    unsigned acc = 0;
    for (int v = 64; v >= 0; --v) {
      const unsigned c = cnt[v];
      cnt[v] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
    const int64_t slot = atomicAdd(&cnt[tile_spread(keys, cl, qtiles, t)], 1u);
    if (slot < nsplit) {  // one of the range's widest: in parts, first in the range
#pragma unroll
      for (int part = 0; part < kSplitParts; ++part)
        order[ilo + kSplitParts * slot + part] = static_cast<int32_t>((t << 3) | (part << 1) | 1);
    } else {
      order[ilo + (kSplitParts - 1) * nsplit + slot] = static_cast<int32_t>(t << 3);
    }
  }
}

// Query side of every call: Morton codes in the cloud boxes, query order, tile prefix,
// tile walk order.
template <typename T>
static hipError_t prep_queries(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  const int bits = key_bits64(L.cl.B);
  const T* q = static_cast<const T*>(L.q);
  const int bs = 256;
  hipLaunchKernelGGL((codes_kernel<T, true>), dim3(static_cast<unsigned>((L.qrows + bs - 1) / bs)), dim3(bs), 0,
                     st, q, L.cl, L.qrows, S.box, S.qkeys);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = radix_sort_pairs<uint64_t>(S.qkeys, S.qkeys_s, nullptr, S.qperm, L.qrows, 0, bits, S.tmp, st);
  if (e != hipSuccess) return e;
  if (L.cl.q_off)
    hipLaunchKernelGGL(tile_prefix_kernel2, dim3(1), dim3(64), 0, st, L.cl.q_off, L.cl.B, int64_t(64), S.qtp);
  if (L.nqtiles > 0)
    hipLaunchKernelGGL(tile_order_kernel, dim3(kXcds), dim3(1024), 0, st, S.qkeys_s, L.cl,
                       L.cl.q_off ? S.qtp : nullptr, L.nqtiles, S.torder);
  return hipGetLastError();
}

template <typename T, int KEY>
static hipError_t prep_pruned(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  if (!L.index) {
    const hipError_t e = prep_points<T, KEY>(L, S, st);
    if (e != hipSuccess) return e;
  }
  if (L.wave) return hipSuccess;  // the per-query walk reads the queries as they are
  return prep_queries<T>(L, S, st);
}

static PrunedState carve_run(const PrunedLaunch& L) {
  size_t tmp = 0;
  (void)sort_temp(L, &tmp);
  return carve(L, tmp);
}

hipError_t launch_knn_pruned_prep(const PrunedLaunch& L, hipStream_t st) {
  const PrunedState S = carve_run(L);
  if (S.total > L.ws_bytes) return hipErrorInvalidValue;
  if (L.dtype == AKNN_F32)
    return L.key == AKNN_KEY_EUCLIDEAN ? prep_pruned<float, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                       : prep_pruned<float, AKNN_KEY_ANISO>(L, S, st);
  return L.key == AKNN_KEY_EUCLIDEAN ? prep_pruned<double, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                     : prep_pruned<double, AKNN_KEY_ANISO>(L, S, st);
}

hipError_t launch_knn_pruned_search(const PrunedLaunch& L, hipStream_t st) {
  const PrunedState S = carve_run(L);
  if (L.wave) {
    if (L.dtype == AKNN_F32)
      return L.key == AKNN_KEY_EUCLIDEAN ? launch_wave<float, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                         : launch_wave<float, AKNN_KEY_ANISO>(L, S, st);
    return L.key == AKNN_KEY_EUCLIDEAN ? launch_wave<double, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                       : launch_wave<double, AKNN_KEY_ANISO>(L, S, st);
  }
  if (L.dtype == AKNN_F32)
    return L.key == AKNN_KEY_EUCLIDEAN ? launch_pruned_topk<float, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                       : launch_pruned_topk<float, AKNN_KEY_ANISO>(L, S, st);
  return L.key == AKNN_KEY_EUCLIDEAN ? launch_pruned_topk<double, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                     : launch_pruned_topk<double, AKNN_KEY_ANISO>(L, S, st);
}

// Workspace of an index build: the point codes and the sort's temporaries.
size_t pruned_index_build_ws(const PrunedLaunch& Lin) {
  PrunedLaunch L = Lin;
  L.qrows = 0;
  L.ways = 1;
  L.ws = nullptr;
  L.index = reinterpret_cast<const void*>(uintptr_t(256));  // only sizes are taken
  size_t tmp = 0;
  (void)sort_temp(L, &tmp);
  return carve(L, tmp).total;
}

// Builds the point index: L.index is the destination, L.ws the temporaries.
hipError_t launch_pruned_index_build(const PrunedLaunch& Lin, hipStream_t st) {
  PrunedLaunch L = Lin;
  L.qrows = 0;
  L.ways = 1;
  const PrunedState S = carve_run(L);
  if (S.total > L.ws_bytes) return hipErrorInvalidValue;
  if (L.dtype == AKNN_F32)
    return L.key == AKNN_KEY_EUCLIDEAN ? prep_points<float, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                       : prep_points<float, AKNN_KEY_ANISO>(L, S, st);
  return L.key == AKNN_KEY_EUCLIDEAN ? prep_points<double, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                     : prep_points<double, AKNN_KEY_ANISO>(L, S, st);
}

const uint32_t* pruned_query_order(const PrunedLaunch& L) { return L.wave ? nullptr : carve_run(L).qperm; }
const uint32_t* pruned_point_order(const PrunedLaunch& L) { return carve_run(L).pperm; }

}  // namespace aknn
