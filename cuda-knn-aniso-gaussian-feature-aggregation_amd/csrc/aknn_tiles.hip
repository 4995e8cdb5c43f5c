// Exact top-K with spatial tile pruning (gfx950).
//
// Same result as the brute-force scan (csrc/forward.cu:83-97 semantics, (key, index)
// order) in far fewer key evaluations.  Per call:
//
//   1. cloud boxes -> 30-bit Morton code per point and per query (10 bits per axis in
//      the cloud's point box); stable radix sorts give a spatial order of both.
//   2. Points are regrouped in that order into tiles of kTile records
//      {p0 p1 p2, R00..R22, s0 s1 s2, original index}; each tile keeps its box and
//      mu_tile = min over its points of a certified lower bound mu_n such that the
//      FP32 key the scan computes satisfies   key(q, n) >= mu_n * |q - p_n|^2.
//   3. One wave = 64 spatially sorted queries.  It visits the cloud's tiles outward
//      from the tile nearest its first query; a tile is scanned only if some lane has
//      mu_tile * dist^2(q, box) <= its current K-th key.  Skipped points provably have
//      keys above every lane's K-th, so the kept set and order are those of a full scan.
//
// mu_n (point with M = R^T diag(s) R): for the exact D = d^T M d >= lmin(M) |d|^2, and
// for the scan's FP32 evaluation (one rounding per op, unit roundoff u, g = 4u/(1-4u)):
//   key >= (1 - g) (D - 2 g |d| sqrt(D S)),   S = sum_i s_i |R_i|^2 = trace(M)
// which is increasing in D where it matters, hence key >= (1-g)(lmin - 2 g sqrt(lmin S)) |d|^2.
// lmin comes from the closed-form 3x3 symmetric eigenvalues in FP64 minus a margin, the
// result is scaled by (1 - 1e-6) to absorb the FP32 rounding of the tile test itself.
// Points with any negative scale, or non-finite data, mark their tile "always scan".
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

constexpr int kTile = 64;
constexpr int kTileSentinel = 0x7fffffff;

// Neighbouring floats (IEEE bit stepping).
__device__ __forceinline__ float next_down(float f) {
  if (f != f || f == -__builtin_inff()) return f;
  if (f == 0.f) return __int_as_float(static_cast<int>(0x80000001u));
  const int b = __float_as_int(f);
  return __int_as_float(f > 0.f ? b - 1 : b + 1);
}
__device__ __forceinline__ float next_up(float f) { return -next_down(-f); }

// ------------------------------------------------------------------ Morton codes
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

template <typename T>
__device__ __forceinline__ uint32_t morton(T x, T y, T z, const float* box) {
  auto q = [](T v, float lo, float hi) -> uint32_t {
    const float ext = hi - lo;
    float t = ext > 0.f ? (static_cast<float>(v) - lo) / ext : 0.f;
    t = t != t ? 0.f : fminf(fmaxf(t, 0.f), 1.f);  // NaN -> 0; clamp queries outside the box
    return static_cast<uint32_t>(t * 1023.f);
  };
  return (spread10(q(x, box[0], box[3])) << 2) | (spread10(q(y, box[1], box[4])) << 1) |
         spread10(q(z, box[2], box[5]));
}

// box[b] = {lo x y z, hi x y z, 0, 0} of the cloud's finite points (float).
template <typename T>
__global__ __launch_bounds__(256) void cloud_box_kernel(const T* __restrict__ p, Clouds cl,
                                                        float* __restrict__ box) {
  const int64_t b = blockIdx.x;
  const int64_t s = cl.p_begin(b), e = cl.p_end(b);
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
  __shared__ float red[6][256];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    red[c][threadIdx.x] = lo[c];
    red[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        red[c][threadIdx.x] = fminf(red[c][threadIdx.x], red[c][threadIdx.x + w]);
        red[3 + c][threadIdx.x] = fmaxf(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) {
    float v = 0.f;
    if (threadIdx.x < 6) v = red[threadIdx.x][0];
    if (threadIdx.x < 6 && !__builtin_isfinite(v)) v = 0.f;  // empty / all non-finite
    box[b * 8 + threadIdx.x] = v;
  }
}

// keys[i] = (cloud << 32) | morton(row i); rows are query rows (QUERY) or point rows.
template <typename T, bool QUERY>
__global__ void codes_kernel(const T* __restrict__ xyz, Clouds cl, int64_t rows,
                             const float* __restrict__ box, uint64_t* __restrict__ keys) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int64_t b = QUERY ? cl.cloud_of_query(i) : cl.cloud_of_point(i);
  const uint32_t c = morton<T>(xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2], box + b * 8);
  keys[i] = (static_cast<uint64_t>(b) << 32) | c;
}

// ------------------------------------------------------------------ point bound
// Certified lower bound mu_n of key(q, n) / |q - p_n|^2 (see the file comment);
// negative means "no bound" (always scan).
template <typename T, int KEY>
__device__ double point_mu(const T* r9, const T* s3) {
  constexpr double u = sizeof(T) == 4 ? 5.9604644775390625e-08 : 1.1102230246251565e-16;
  constexpr double g = 4.0 * u / (1.0 - 4.0 * u);
  if (KEY == AKNN_KEY_EUCLIDEAN) return (1.0 - g) * (1.0 - 2.0 * g * 1.7320508075688772);
  double R[9], s[3];
  for (int i = 0; i < 9; ++i) R[i] = static_cast<double>(r9[i]);
  for (int i = 0; i < 3; ++i) s[i] = static_cast<double>(s3[i]);
  for (int i = 0; i < 9; ++i)
    if (!__builtin_isfinite(R[i])) return -1.0;
  for (int i = 0; i < 3; ++i)
    if (!(s[i] >= 0.0) || !__builtin_isfinite(s[i])) return -1.0;
  // M = R^T diag(s) R (symmetric 3x3)
  double M[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c)
      M[a][c] = s[0] * R[a] * R[c] + s[1] * R[3 + a] * R[3 + c] + s[2] * R[6 + a] * R[6 + c];
  const double S = M[0][0] + M[1][1] + M[2][2];
  const double m = S / 3.0;
  const double p1 = M[0][1] * M[0][1] + M[0][2] * M[0][2] + M[1][2] * M[1][2];
  const double a0 = M[0][0] - m, a1 = M[1][1] - m, a2 = M[2][2] - m;
  const double p2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * p1;
  double lmin;
  if (p2 <= 0.0) {
    lmin = m;
  } else {
    const double pp = sqrt(p2 / 6.0);
    const double b00 = a0 / pp, b11 = a1 / pp, b22 = a2 / pp;
    const double b01 = M[0][1] / pp, b02 = M[0][2] / pp, b12 = M[1][2] / pp;
    const double det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                       b02 * (b01 * b12 - b11 * b02);
    double rr = det / 2.0;
    rr = rr < -1.0 ? -1.0 : (rr > 1.0 ? 1.0 : rr);
    const double phi = acos(rr) / 3.0;
    lmin = m + 2.0 * pp * cos(phi + 2.0943951023931957);  // smallest eigenvalue
  }
  // closed-form eigenvalues are accurate to a few ulp of the largest one
  lmin -= 1e-12 * (fabs(m) + S);
  if (!(lmin > 0.0)) return 0.0;
  const double mu = (1.0 - g) * (lmin - 2.0 * g * sqrt(lmin * S));
  return mu > 0.0 ? mu : 0.0;
}

// ------------------------------------------------------------------ tile build
// One wave per tile.  rec: [tiles*kTile][16] records; hdr: [tiles][8] =
// {lo x y z, hi x y z, mu (float, rounded down; < 0 = always scan), first code (bits)}.
template <typename T, int KEY>
__global__ __launch_bounds__(64) void tile_build_kernel(
    const T* __restrict__ p, const T* __restrict__ sigma, const T* __restrict__ R, Clouds cl,
    const int64_t* __restrict__ tile_prefix, const uint32_t* __restrict__ perm,
    const uint64_t* __restrict__ sorted_keys, T* __restrict__ rec, float* __restrict__ hdr,
    int32_t* __restrict__ tile_cnt) {
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
  T* out = rec + (t * kTile + lane) * 16;
  if (lane < cnt) {
    const int64_t row = perm[i0 + lane];
    T v[16];
    for (int c = 0; c < 3; ++c) v[c] = p[row * 3 + c];
    for (int c = 0; c < 9; ++c) v[3 + c] = R[row * 9 + c];
    for (int c = 0; c < 3; ++c) v[12 + c] = sigma[row * 3 + c];
    const int idx = static_cast<int>(row - pb);
    if constexpr (sizeof(T) == 4) v[15] = __int_as_float(idx);
    else v[15] = __longlong_as_double(static_cast<long long>(idx));
    for (int c = 0; c < 16; ++c) out[c] = v[c];
    for (int c = 0; c < 3; ++c) {
      // box in float, widened outward so it contains the exact coordinate
      const float f = static_cast<float>(v[c]);
      if (!__builtin_isfinite(f)) bad = true;
      lo3[c] = static_cast<double>(f) <= static_cast<double>(v[c]) ? f : next_down(f);
      hi3[c] = static_cast<double>(f) >= static_cast<double>(v[c]) ? f : next_up(f);
    }
    mu = point_mu<T, KEY>(v + 3, v + 12);
    if (!(mu >= 0.0)) bad = true;
  } else {
    for (int c = 0; c < 16; ++c) out[c] = T(0);
  }
  // wave reductions (64 lanes)
  for (int off = 32; off > 0; off >>= 1) {
    for (int c = 0; c < 3; ++c) {
      lo3[c] = fminf(lo3[c], __shfl_xor(lo3[c], off, 64));
      hi3[c] = fmaxf(hi3[c], __shfl_xor(hi3[c], off, 64));
    }
    mu = fmin(mu, __shfl_xor(mu, off, 64));
    bad = bad | (__shfl_xor(static_cast<int>(bad), off, 64) != 0);
  }
  if (lane == 0) {
    float* h = hdr + t * 8;
    for (int c = 0; c < 3; ++c) {
      h[c] = lo3[c];
      h[3 + c] = hi3[c];
    }
    float muf = -1.f;
    if (!bad) {
      const double m = mu * (1.0 - 1e-6);
      muf = static_cast<float>(m);
      if (static_cast<double>(muf) > m) muf = next_down(muf);
    }
    h[6] = muf;
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

template <typename T>
__device__ __forceinline__ int rec_index(T v) {
  if constexpr (sizeof(T) == 4) return __float_as_int(v);
  else return static_cast<int>(__double_as_longlong(v));
}

// grid.x = query tiles of 64 sorted queries (ragged: via q tile prefix), grid.y = clouds.
template <typename T, int KMAX, int KEY>
__global__ __launch_bounds__(64) void knn_pruned_kernel(
    const T* __restrict__ q, Clouds cl, const int64_t* __restrict__ qtiles,
    const uint32_t* __restrict__ qperm, const uint64_t* __restrict__ qkeys,
    const int64_t* __restrict__ ptiles, const T* __restrict__ rec, const float* __restrict__ hdr,
    const int32_t* __restrict__ tcnt, int K, T* __restrict__ dists, int64_t* __restrict__ idxs) {
  const int lane = threadIdx.x;
  int64_t b, q0;
  if (qtiles) {
    const int64_t t = blockIdx.x;
    int64_t lo = 0, hi = cl.B - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (qtiles[mid] <= t) lo = mid; else hi = mid - 1;
    }
    b = lo;
    q0 = cl.q_begin(b) + (t - qtiles[b]) * 64;
  } else {
    b = blockIdx.y;
    q0 = cl.q_begin(b) + static_cast<int64_t>(blockIdx.x) * 64;
  }
  const int64_t qend = cl.q_end(b);
  const int64_t pos = q0 + lane;
  const bool valid = pos < qend;
  int64_t row = 0;
  T qx = __builtin_nan(""), qy = qx, qz = qx;
  if (valid) {
    row = qperm[pos];
    qx = q[row * 3 + 0];
    qy = q[row * 3 + 1];
    qz = q[row * 3 + 2];
  }

  T key[KMAX];
  int id[KMAX];
#pragma unroll
  for (int s = 0; s < KMAX; ++s) {
    key[s] = __builtin_inf();
    id[s] = kTileSentinel;
  }

  // this cloud's tiles [tb, te)
  int64_t tb, te;
  if (cl.p_off) {
    tb = ptiles[b];
    te = ptiles[b + 1];
  } else {
    tb = b * ptiles[1];
    te = tb + ptiles[1];
  }
  // seed: last tile whose first code <= the wave's first query code (scalar search)
  const uint32_t code0 = static_cast<uint32_t>(qkeys[q0] & 0xffffffffu);
  int64_t lo = tb, hi = te - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (__float_as_uint(hdr[mid * 8 + 7]) <= code0) lo = mid; else hi = mid - 1;
  }
  const int64_t seed = __builtin_amdgcn_readfirstlane(static_cast<int>(lo - tb)) + tb;
  const int64_t ntiles = te - tb;

  // outward walk: seed, seed+1, seed-1, seed+2, ...
  int64_t up = seed, down = seed - 1;
  for (int64_t step = 0; step < ntiles; ++step) {
    int64_t t;
    const bool take_up = (up < te) && ((step & 1) == 0 || down < tb);
    if (take_up) t = up++;
    else t = down--;
    // lower bound of every key in the tile, in the key's precision: the box holds the
    // exact coordinates, so each difference is rounded relative to itself
    const float* h = hdr + t * 8;
    const T mu = static_cast<T>(h[6]);
    const T dx = fmax(fmax(static_cast<T>(h[0]) - qx, qx - static_cast<T>(h[3])), T(0));
    const T dy = fmax(fmax(static_cast<T>(h[1]) - qy, qy - static_cast<T>(h[4])), T(0));
    const T dz = fmax(fmax(static_cast<T>(h[2]) - qz, qz - static_cast<T>(h[5])), T(0));
    const T lb = mu * (dx * dx + dy * dy + dz * dz);
    const bool need = valid & ((mu < T(0)) | !(lb > key[KMAX - 1]));
    if (!__any(need)) continue;
    const int cnt = tcnt[t];
    const T* __restrict__ base = rec + t * (kTile * 16);
    for (int j = 0; j < cnt; ++j) {
      const T* r = base + j * 16;
      const T c = pair_key<T, KEY>(qx, qy, qz, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7],
                                   r[8], r[9], r[10], r[11], r[12], r[13], r[14]);
      if (c <= key[KMAX - 1]) lex_insert<KMAX>(key, id, c, rec_index(r[15]));
    }
  }

  if (valid) {
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      if (s < K) {
        dists[row * K + s] = key[s];
        idxs[row * K + s] = static_cast<int64_t>(id[s]);
      }
    }
  }
}

template <typename T, int KMAX, int KEY>
static hipError_t launch_pruned_k(const PrunedLaunch& L, const uint64_t* qkeys, const uint32_t* qperm,
                                  const int64_t* ptp, const int64_t* qtp, const T* rec,
                                  const float* hdr, const int32_t* tcnt, hipStream_t st) {
  dim3 grid;
  if (L.cl.q_off) grid = dim3(static_cast<unsigned>(L.nqtiles), 1);
  else grid = dim3(static_cast<unsigned>((L.cl.Q + 63) / 64), static_cast<unsigned>(L.cl.B));
  hipLaunchKernelGGL((knn_pruned_kernel<T, KMAX, KEY>), grid, dim3(64), 0, st, static_cast<const T*>(L.q),
                     L.cl, L.cl.q_off ? qtp : nullptr, qperm, qkeys, ptp, rec, hdr, tcnt,
                     static_cast<int>(L.K), static_cast<T*>(L.dists), L.idxs);
  return hipGetLastError();
}

template <typename T, int KEY>
static hipError_t launch_pruned_topk(const PrunedLaunch& L, const uint64_t* qkeys, const uint32_t* qperm,
                                     const int64_t* ptp, const int64_t* qtp, const T* rec,
                                     const float* hdr, const int32_t* tcnt, hipStream_t st) {
  if (L.K <= 1) return launch_pruned_k<T, 1, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
  if (L.K <= 4) return launch_pruned_k<T, 4, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
  if (L.K <= 8) return launch_pruned_k<T, 8, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
  if (L.K <= 16) return launch_pruned_k<T, 16, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
  if (L.K <= 32) return launch_pruned_k<T, 32, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
  return launch_pruned_k<T, 64, KEY>(L, qkeys, qperm, ptp, qtp, rec, hdr, tcnt, st);
}

// ------------------------------------------------------------------ host side
static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

static int key_bits64(int64_t B) {
  int bits = 32;
  while ((int64_t(1) << (bits - 32)) < B) ++bits;
  return bits;
}

static hipError_t sort_bytes(int64_t n, int bits, size_t* out) {
  rocprim::counting_iterator<uint32_t> vin(0);
  return rocprim::radix_sort_pairs(nullptr, *out, static_cast<const uint64_t*>(nullptr),
                                   static_cast<uint64_t*>(nullptr), vin,
                                   static_cast<uint32_t*>(nullptr), static_cast<size_t>(n), 0, bits,
                                   nullptr);
}

// Workspace layout (each piece 256-B aligned):
//   box[B*8] f32 | pkeys, pkeys_sorted [Prows] u64 | pperm [Prows] u32 |
//   qkeys, qkeys_sorted [Qrows] u64 | qperm [Qrows] u32 | ptile_prefix, qtile_prefix [B+1] i64 |
//   rec [T*64*16] | hdr [T*8] f32 | tcnt [T] i32 | sort temp (max of both sorts)
hipError_t pruned_workspace(const PrunedLaunch& L, size_t* bytes) {
  const int bits = key_bits64(L.cl.B);
  size_t sp = 0, sq = 0;
  hipError_t e = sort_bytes(L.prows, bits, &sp);
  if (e != hipSuccess) return e;
  e = sort_bytes(L.qrows, bits, &sq);
  if (e != hipSuccess) return e;
  const size_t el = L.dtype == AKNN_F64 ? 8 : 4;
  *bytes = al(L.cl.B * 8 * 4) + 2 * al(L.prows * 8) + al(L.prows * 4) + 2 * al(L.qrows * 8) +
           al(L.qrows * 4) + 2 * al((L.cl.B + 1) * 8) + al(L.ntiles * kTile * 16 * el) +
           al(L.ntiles * 8 * 4) + al(L.ntiles * 4) + al(sp > sq ? sp : sq);
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

__global__ void dense_prefix_kernel(int64_t per_cloud_tiles, int64_t* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = 0;
    out[1] = per_cloud_tiles;
  }
}

// Carve the workspace (layout above).
struct PrunedState {
  float* box;
  uint64_t *pkeys, *pkeys_s, *qkeys, *qkeys_s;
  uint32_t *pperm, *qperm;
  int64_t *ptp, *qtp;
  void* rec;
  float* hdr;
  int32_t* tcnt;
  void* tmp;
  size_t tmp_bytes;
};

static PrunedState carve(const PrunedLaunch& L) {
  PrunedState S{};
  char* w = static_cast<char*>(L.ws);
  auto take = [&](size_t n) {
    char* p = w;
    w += al(n);
    return p;
  };
  const size_t el = L.dtype == AKNN_F64 ? 8 : 4;
  S.box = reinterpret_cast<float*>(take(L.cl.B * 8 * 4));
  S.pkeys = reinterpret_cast<uint64_t*>(take(L.prows * 8));
  S.pkeys_s = reinterpret_cast<uint64_t*>(take(L.prows * 8));
  S.pperm = reinterpret_cast<uint32_t*>(take(L.prows * 4));
  S.qkeys = reinterpret_cast<uint64_t*>(take(L.qrows * 8));
  S.qkeys_s = reinterpret_cast<uint64_t*>(take(L.qrows * 8));
  S.qperm = reinterpret_cast<uint32_t*>(take(L.qrows * 4));
  S.ptp = reinterpret_cast<int64_t*>(take((L.cl.B + 1) * 8));
  S.qtp = reinterpret_cast<int64_t*>(take((L.cl.B + 1) * 8));
  S.rec = take(L.ntiles * kTile * 16 * el);
  S.hdr = reinterpret_cast<float*>(take(L.ntiles * 8 * 4));
  S.tcnt = reinterpret_cast<int32_t*>(take(L.ntiles * 4));
  S.tmp = w;
  S.tmp_bytes = L.ws_bytes - static_cast<size_t>(w - static_cast<char*>(L.ws));
  return S;
}

// Boxes, Morton codes, both sorts and the tile records/headers.
template <typename T, int KEY>
static hipError_t prep_pruned(const PrunedLaunch& L, const PrunedState& S, hipStream_t st) {
  const int bits = key_bits64(L.cl.B);
  const T* q = static_cast<const T*>(L.q);
  const T* p = static_cast<const T*>(L.p);
  hipLaunchKernelGGL((cloud_box_kernel<T>), dim3(static_cast<unsigned>(L.cl.B)), dim3(256), 0, st, p, L.cl, S.box);
  const int bs = 256;
  hipLaunchKernelGGL((codes_kernel<T, false>), dim3(static_cast<unsigned>((L.prows + bs - 1) / bs)), dim3(bs), 0,
                     st, p, L.cl, L.prows, S.box, S.pkeys);
  hipLaunchKernelGGL((codes_kernel<T, true>), dim3(static_cast<unsigned>((L.qrows + bs - 1) / bs)), dim3(bs), 0,
                     st, q, L.cl, L.qrows, S.box, S.qkeys);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  rocprim::counting_iterator<uint32_t> vin(0);
  size_t tb = S.tmp_bytes;
  e = rocprim::radix_sort_pairs(S.tmp, tb, static_cast<const uint64_t*>(S.pkeys), S.pkeys_s, vin, S.pperm,
                                static_cast<size_t>(L.prows), 0, bits, st);
  if (e != hipSuccess) return e;
  tb = S.tmp_bytes;
  e = rocprim::radix_sort_pairs(S.tmp, tb, static_cast<const uint64_t*>(S.qkeys), S.qkeys_s, vin, S.qperm,
                                static_cast<size_t>(L.qrows), 0, bits, st);
  if (e != hipSuccess) return e;
  if (L.cl.p_off) {
    hipLaunchKernelGGL(tile_prefix_kernel2, dim3(1), dim3(64), 0, st, L.cl.p_off, L.cl.B, int64_t(kTile), S.ptp);
    hipLaunchKernelGGL(tile_prefix_kernel2, dim3(1), dim3(64), 0, st, L.cl.q_off, L.cl.B, int64_t(64), S.qtp);
  } else {
    hipLaunchKernelGGL(dense_prefix_kernel, dim3(1), dim3(64), 0, st, (L.cl.P + kTile - 1) / kTile, S.ptp);
  }
  hipLaunchKernelGGL((tile_build_kernel<T, KEY>), dim3(static_cast<unsigned>(L.ntiles)), dim3(64), 0, st,
                     p, static_cast<const T*>(L.sigma), static_cast<const T*>(L.R), L.cl, S.ptp, S.pperm,
                     S.pkeys_s, static_cast<T*>(S.rec), S.hdr, S.tcnt);
  return hipGetLastError();
}

hipError_t launch_knn_pruned_prep(const PrunedLaunch& L, hipStream_t st) {
  const PrunedState S = carve(L);
  if (L.dtype == AKNN_F32)
    return L.key == AKNN_KEY_EUCLIDEAN ? prep_pruned<float, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                       : prep_pruned<float, AKNN_KEY_ANISO>(L, S, st);
  return L.key == AKNN_KEY_EUCLIDEAN ? prep_pruned<double, AKNN_KEY_EUCLIDEAN>(L, S, st)
                                     : prep_pruned<double, AKNN_KEY_ANISO>(L, S, st);
}

hipError_t launch_knn_pruned_search(const PrunedLaunch& L, hipStream_t st) {
  const PrunedState S = carve(L);
  if (L.dtype == AKNN_F32) {
    const float* rec = static_cast<const float*>(S.rec);
    return L.key == AKNN_KEY_EUCLIDEAN
               ? launch_pruned_topk<float, AKNN_KEY_EUCLIDEAN>(L, S.qkeys_s, S.qperm, S.ptp, S.qtp, rec, S.hdr, S.tcnt, st)
               : launch_pruned_topk<float, AKNN_KEY_ANISO>(L, S.qkeys_s, S.qperm, S.ptp, S.qtp, rec, S.hdr, S.tcnt, st);
  }
  const double* rec = static_cast<const double*>(S.rec);
  return L.key == AKNN_KEY_EUCLIDEAN
             ? launch_pruned_topk<double, AKNN_KEY_EUCLIDEAN>(L, S.qkeys_s, S.qperm, S.ptp, S.qtp, rec, S.hdr, S.tcnt, st)
             : launch_pruned_topk<double, AKNN_KEY_ANISO>(L, S.qkeys_s, S.qperm, S.ptp, S.qtp, rec, S.hdr, S.tcnt, st);
}

const uint32_t* pruned_query_order(const PrunedLaunch& L) { return carve(L).qperm; }

}  // namespace aknn
