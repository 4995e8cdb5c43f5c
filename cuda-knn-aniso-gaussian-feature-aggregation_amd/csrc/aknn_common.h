// Shared device helpers for the gfx950 knn_aggregate_aniso kernels.
//
// Geometry per (query q, point n), with d = q - p_n:
//   u = R_n d                      (csrc/forward.cu:88-90)
//   D = s0 u0^2 + s1 u1^2 + s2 u2^2 (csrc/forward.cu:92)
//   w = exp(-D)                    (csrc/forward.cu:102)
//   v = R_n^T (s_n * u)            (csrc/backward.cu:207-209)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/aknn.h"

namespace aknn {

constexpr int kWave = 64;
// Lanes that cooperate on one query row (or one point row) in the gather kernels:
// one DPP row, so the cross-lane steps stay inside a row.
constexpr int kGroup = 16;

// ------------------------------------------------------------------ row addressing
// Dense batches: cloud b owns query rows [b*Q, (b+1)*Q) and point rows [b*P, (b+1)*P).
// Ragged batches: offsets arrays of B+1 int64 on the device.
struct Clouds {
  const int64_t* q_off;
  const int64_t* p_off;
  int64_t B, Q, P;  // dense per-cloud sizes, or totals for ragged batches
  // The row arguments below are rows of the batch: [0, B*Q) queries, [0, B*P) points.

  // Cloud of a dense row (row < B * n): no division for one cloud, a 32-bit one when the
  // operands fit (the 64-bit division is a long scalar and vector sequence).
  __device__ __forceinline__ int64_t dense_cloud(int64_t row, int64_t n) const {
    if (B == 1) return 0;
    if (((static_cast<uint64_t>(row) | static_cast<uint64_t>(n)) >> 32) == 0)
      return static_cast<int64_t>(static_cast<uint32_t>(row) / static_cast<uint32_t>(n));
    return row / n;
  }
  __device__ __forceinline__ int64_t cloud_of_query(int64_t row) const {
    if (!q_off) return dense_cloud(row, Q);
    int64_t lo = 0, hi = B - 1;  // last b with q_off[b] <= row
    while (lo < hi) {
      int64_t mid = (lo + hi + 1) >> 1;
      if (q_off[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
  }
  __device__ __forceinline__ int64_t cloud_of_point(int64_t row) const {
    if (!p_off) return dense_cloud(row, P);
    int64_t lo = 0, hi = B - 1;  // last b with p_off[b] <= row
    while (lo < hi) {
      int64_t mid = (lo + hi + 1) >> 1;
      if (p_off[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
  }
  __device__ __forceinline__ int64_t q_begin(int64_t b) const { return q_off ? q_off[b] : b * Q; }
  __device__ __forceinline__ int64_t q_end(int64_t b) const { return q_off ? q_off[b + 1] : (b + 1) * Q; }
  __device__ __forceinline__ int64_t p_begin(int64_t b) const { return p_off ? p_off[b] : b * P; }
  __device__ __forceinline__ int64_t p_end(int64_t b) const { return p_off ? p_off[b + 1] : (b + 1) * P; }
};

// ------------------------------------------------------------------ XCD-aware block order
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with its
// own L2.  Kernels whose neighbouring logical blocks share data (spatially ordered
// rows) remap the hardware block id so each XCD works through one contiguous range of
// logical blocks.  A bijection on [0, n) for any n; results never depend on it.
constexpr int kXcds = 8;

// Point index of an unfilled top-K slot inside the search kernels.
constexpr int kNoPoint = 0x7fffffff;

// A query with fewer than K keys that compare (NaN query coordinates, or a cloud with
// fewer than K points whose keys are not NaN) leaves the tail of its list unfilled.
// Those slots take the lowest point indices not already in the list, with a NaN key:
// every index stays inside the cloud, and a NaN query gets 0..K-1 with NaN keys, which
// is what the reference's scan keeps when no key compares (csrc/mink.cuh:60-81 fills
// its K slots with the first points scanned and never replaces them).  Called, rarely,
// after the row's slots have been stored by the same thread.
template <typename T>
__device__ __noinline__ void fill_unfilled_row(T* __restrict__ d, int64_t* __restrict__ ix, int K) {
  int64_t c = 0;
  for (int s = 0; s < K; ++s) {
    if (ix[s] != kNoPoint) continue;
    for (;;) {
      bool used = false;
      for (int t = 0; t < K; ++t) used |= ix[t] == c;
      if (!used) break;
      ++c;
    }
    ix[s] = c++;
    d[s] = static_cast<T>(__builtin_nan(""));
  }
}
__device__ __forceinline__ int64_t xcd_block(int64_t hw, int64_t n) {
  const int64_t q = n / kXcds, r = n % kXcds;
  const int64_t x = hw % kXcds, i = hw / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// Exact unsigned 32-bit division by a divisor fixed for the whole kernel (the neighbour
// count K of an entry number q*K + k), by multiply-high and shifts: with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1 and t = mulhi(m, n), n / d = (t + ((n - t) >> min(l,
// 1))) >> max(l - 1, 0) for every 32-bit n (Granlund and Montgomery, "Division by
// invariant integers using multiplication", 1994, section 4).
struct FastDiv {
  uint32_t m;
  int s1, s2;
  __device__ __forceinline__ explicit FastDiv(uint32_t d) {
    int l = 0;
    while ((1u << l) < d) ++l;
    m = static_cast<uint32_t>(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    s1 = l < 1 ? l : 1;
    s2 = l > 1 ? l - 1 : 0;
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> s1)) >> s2;
  }
};

// ------------------------------------------------------------------ key
// The neighbour key, one rounding per operation in the reference's association
// order (csrc/forward.cu:85-92 and :154).  Contraction into FMA is disabled here so
// the result is the same bits as the CPU formulation; everything else in the
// library may contract.
template <typename T, int KEY>
__device__ __forceinline__ T pair_key(T qx, T qy, T qz, T px, T py, T pz, T r00, T r01, T r02,
                                      T r10, T r11, T r12, T r20, T r21, T r22, T s0, T s1,
                                      T s2) {
#pragma clang fp contract(off)
  T x = qx - px;
  T y = qy - py;
  T z = qz - pz;
  if (KEY == AKNN_KEY_EUCLIDEAN) {
    T e = x * x;
    e = e + y * y;
    e = e + z * z;
    return e;
  }
  T j = r00 * x;
  j = j + r01 * y;
  j = j + r02 * z;
  T k = r10 * x;
  k = k + r11 * y;
  k = k + r12 * z;
  T l = r20 * x;
  l = l + r21 * y;
  l = l + r22 * z;
  T D = (j * j) * s0;
  D = D + (k * k) * s1;
  D = D + (l * l) * s2;
  return D;
}

__device__ __forceinline__ float gexp(float x) { return expf(x); }
__device__ __forceinline__ double gexp(double x) { return exp(x); }

// Per-(query, point) quantities recomputed by the backward kernels (never stored).
template <typename T>
struct Geo {
  T d0, d1, d2;  // d = q - p
  T u0, u1, u2;  // u = R d
  T w;           // exp(-D), D by pair_key (same bits as the forward)
  T v0, v1, v2;  // R^T (s * u)
};

// Point record as loaded by the backward kernels.
template <typename T>
struct Pt {
  T p0, p1, p2;
  T r[9];
  T s0, s1, s2;
};

template <typename T>
__device__ __forceinline__ Pt<T> load_point(const T* __restrict__ p, const T* __restrict__ sigma,
                                            const T* __restrict__ R, int64_t row) {
  Pt<T> o;
  o.p0 = p[row * 3 + 0];
  o.p1 = p[row * 3 + 1];
  o.p2 = p[row * 3 + 2];
#pragma unroll
  for (int i = 0; i < 9; ++i) o.r[i] = R[row * 9 + i];
  o.s0 = sigma[row * 3 + 0];
  o.s1 = sigma[row * 3 + 1];
  o.s2 = sigma[row * 3 + 2];
  return o;
}

// Point record from the packed [rows][16] array {p0 p1 p2, R00..R22, s0 s1 s2, -}:
// 64-B (FP32) or 128-B (FP64) aligned, so four or eight 16-B loads.
template <typename T>
__device__ __forceinline__ Pt<T> load_record(const T* __restrict__ rec, int64_t row) {
  Pt<T> o;
  const T* r = rec + row * 16;
  T v[16];
  if constexpr (sizeof(T) == 4) {
    const float4* r4 = reinterpret_cast<const float4*>(r);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 x = r4[i];
      v[4 * i] = x.x;
      v[4 * i + 1] = x.y;
      v[4 * i + 2] = x.z;
      v[4 * i + 3] = x.w;
    }
  } else {
    const double2* r2 = reinterpret_cast<const double2*>(r);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double2 x = r2[i];
      v[2 * i] = x.x;
      v[2 * i + 1] = x.y;
    }
  }
  o.p0 = v[0];
  o.p1 = v[1];
  o.p2 = v[2];
#pragma unroll
  for (int i = 0; i < 9; ++i) o.r[i] = v[3 + i];
  o.s0 = v[12];
  o.s1 = v[13];
  o.s2 = v[14];
  return o;
}

template <typename T>
__device__ __forceinline__ Pt<T> fetch_point(const T* __restrict__ rec, const T* __restrict__ p,
                                             const T* __restrict__ sigma, const T* __restrict__ R,
                                             int64_t row) {
  return rec ? load_record<T>(rec, row) : load_point<T>(p, sigma, R, row);
}

// Weights are always anisotropic; the Euclidean key mode only changes the ranking
// (as in the reference's lookup path, csrc/forward.cu:165-181).
template <typename T>
__device__ __forceinline__ Geo<T> geometry(T qx, T qy, T qz, const Pt<T>& P) {
  Geo<T> g;
  const T D = pair_key<T, AKNN_KEY_ANISO>(qx, qy, qz, P.p0, P.p1, P.p2, P.r[0], P.r[1], P.r[2], P.r[3],
                               P.r[4], P.r[5], P.r[6], P.r[7], P.r[8], P.s0, P.s1, P.s2);
  g.w = gexp(-D);
  g.d0 = qx - P.p0;
  g.d1 = qy - P.p1;
  g.d2 = qz - P.p2;
  g.u0 = P.r[0] * g.d0 + P.r[1] * g.d1 + P.r[2] * g.d2;
  g.u1 = P.r[3] * g.d0 + P.r[4] * g.d1 + P.r[5] * g.d2;
  g.u2 = P.r[6] * g.d0 + P.r[7] * g.d1 + P.r[8] * g.d2;
  const T a0 = P.s0 * g.u0, a1 = P.s1 * g.u1, a2 = P.s2 * g.u2;
  g.v0 = P.r[0] * a0 + P.r[3] * a1 + P.r[6] * a2;
  g.v1 = P.r[1] * a0 + P.r[4] * a1 + P.r[7] * a2;
  g.v2 = P.r[2] * a0 + P.r[5] * a1 + P.r[8] * a2;
  return g;
}

// ------------------------------------------------------------------ full second order
// Upstream gradients of one point's first-order outputs (Hp = d/d dp, Hs = d/d dsigma,
// HR = d/d dR, row-major); zero when absent.
template <typename T>
struct PtUp {
  T p[3], s[3], r[9];
};

template <typename T>
__device__ __forceinline__ PtUp<T> load_up(const T* __restrict__ gp, const T* __restrict__ gs,
                                           const T* __restrict__ gr, int64_t row) {
  PtUp<T> u;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    u.p[i] = gp ? gp[row * 3 + i] : T(0);
    u.s[i] = gs ? gs[row * 3 + i] : T(0);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) u.r[i] = gr ? gr[row * 9 + i] : T(0);
  return u;
}

// Per (query, point) entry of the second order for upstream (h = grad_grad_q[q], Hp,
// Hs, HR): the entry adds  c_e = a w Phi + w (g . Hf)  to <upstream, first-order
// outputs>, with hh = h - Hp, s = v . hh and
//   Phi = -2 s - sum_i Hs_i u_i^2 - 2 sum_i sig_i u_i (HR_i . d).
// With only h this is the reference's -2 s (csrc/backward.cu:196-209).
template <typename T>
struct Second2 {
  T hh[3];   // h - Hp
  T rh[3];   // R_i . hh
  T hrd[3];  // HR_i . d
  T phi;
};

template <typename T>
__device__ __forceinline__ Second2<T> second_phi(const Geo<T>& g, const Pt<T>& P, T hx, T hy, T hz,
                                                 const PtUp<T>& up) {
  Second2<T> o;
  o.hh[0] = hx - up.p[0];
  o.hh[1] = hy - up.p[1];
  o.hh[2] = hz - up.p[2];
  const T u[3] = {g.u0, g.u1, g.u2};
  const T dd[3] = {g.d0, g.d1, g.d2};
  const T sg[3] = {P.s0, P.s1, P.s2};
  T phi = T(-2) * (o.hh[0] * g.v0 + o.hh[1] * g.v1 + o.hh[2] * g.v2);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    o.rh[i] = P.r[3 * i] * o.hh[0] + P.r[3 * i + 1] * o.hh[1] + P.r[3 * i + 2] * o.hh[2];
    o.hrd[i] = up.r[3 * i] * dd[0] + up.r[3 * i + 1] * dd[1] + up.r[3 * i + 2] * dd[2];
    phi -= up.s[i] * u[i] * u[i] + T(2) * sg[i] * u[i] * o.hrd[i];
  }
  o.phi = phi;
  return o;
}

// ------------------------------------------------------------------ cross-lane
// DPP moves inside a row of 16 lanes (gfx9 encodings).
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              (static_cast<long long>(static_cast<unsigned>(lo))));
}
template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

// Partner of each lane in the 4 butterfly stages (bit 3, 2, 1, 0 of the lane in its
// row): mirror pairs l <-> 15-l, half mirror l <-> 7-l, then quad xor 2 and xor 1.
// Each pairing joins lanes that agree on the higher bits and differ in the stage bit.
template <int STAGE, typename T>
__device__ __forceinline__ T bfly(T v) {
  if (STAGE == 3) return dpp<kDppMirror>(v);
  if (STAGE == 2) return dpp<kDppHalfMirror>(v);
  if (STAGE == 1) return dpp<kDppQuadXor2>(v);
  return dpp<kDppQuadXor1>(v);
}

// Sum of v over the 16 lanes of the row, in every lane.  Fixed tree: deterministic.
template <typename T>
__device__ __forceinline__ T row_sum(T v) {
  v = v + bfly<3>(v);
  v = v + bfly<2>(v);
  v = v + bfly<1>(v);
  v = v + bfly<0>(v);
  return v;
}

// Reduce-scatter of 16 partial sums: on entry lane r of a row holds x[0..15], one
// partial per entry j; on exit lane r returns sum over the row's lanes of x[r].
// 15 DPP moves; fixed order, so the result is deterministic.
template <typename T>
__device__ __forceinline__ T row_transpose_sum(T (&x)[16], int lane_in_row) {
  // stage bit 3: keep entries whose bit 3 equals the lane's bit 3
  const bool b3 = (lane_in_row >> 3) & 1;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const T mine = b3 ? x[8 + i] : x[i];
    const T give = b3 ? x[i] : x[8 + i];
    x[i] = mine + bfly<3>(give);
  }
  const bool b2 = (lane_in_row >> 2) & 1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const T mine = b2 ? x[4 + i] : x[i];
    const T give = b2 ? x[i] : x[4 + i];
    x[i] = mine + bfly<2>(give);
  }
  const bool b1 = (lane_in_row >> 1) & 1;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const T mine = b1 ? x[2 + i] : x[i];
    const T give = b1 ? x[i] : x[2 + i];
    x[i] = mine + bfly<1>(give);
  }
  const bool b0 = lane_in_row & 1;
  const T mine = b0 ? x[1] : x[0];
  const T give = b0 ? x[0] : x[1];
  return mine + bfly<0>(give);
}

// Value of v held by lane `src` (0..15) of this lane's row.
template <typename T>
__device__ __forceinline__ T row_bcast(T v, int src) {
  const int lane = __lane_id();
  return __shfl(v, (lane & ~15) | src, 64);
}

}  // namespace aknn
