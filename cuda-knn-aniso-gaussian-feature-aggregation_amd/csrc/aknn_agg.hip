// Query-side gather kernels for gfx950: forward aggregation, first-order dq and
// second-order (dq, d grad_f_out).
//
// Replaces the per-query loops of
//   knn_aggregate_aniso_forward_cuda_kernel      csrc/forward.cu:99-109
//   knn_aggregate_aniso_backward_cuda_kernel     csrc/backward.cu:31-98  (dq part)
//   knn_aggregate_aniso_backward_2nd_cuda_kernel csrc/backward.cu:156-305 (dq, dgrad_f_out)
//
// Mapping: one row of 16 lanes per query (4 queries per wave).  Lane r owns
// neighbour slot k = kb + r of each batch of 16 slots: it loads that slot's index,
// recomputes the slot's geometry (never read back from memory) and, after the
// cross-lane reduce-scatter, holds that slot's a_k = grad_f_out[q] . f[n_k].
// Features are spread over the row (VEC consecutive values per lane and chunk), so
// every f row gather is one contiguous, coalesced read of F values.
#include <type_traits>

#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

template <typename T, int VEC>
struct alignas(sizeof(T) * VEC) VecT {
  T v[VEC];
};

template <typename T, int VEC>
__device__ __forceinline__ void vload(T (&d)[VEC], const T* __restrict__ src) {
  const VecT<T, VEC> x = *reinterpret_cast<const VecT<T, VEC>*>(src);
#pragma unroll
  for (int i = 0; i < VEC; ++i) d[i] = x.v[i];
}

template <typename T, int VEC>
__device__ __forceinline__ void vstore(T* __restrict__ dst, const T (&s)[VEC]) {
  VecT<T, VEC> x;
#pragma unroll
  for (int i = 0; i < VEC; ++i) x.v[i] = s[i];
  *reinterpret_cast<VecT<T, VEC>*>(dst) = x;
}

// Streaming store (non-temporal): outputs that nothing in the kernel reads again, so
// they do not push the gathered point and feature rows out of L2.
template <typename T, int VEC>
__device__ __forceinline__ void vstore_nt(T* __restrict__ dst, const T (&s)[VEC]) {
  typedef T V __attribute__((ext_vector_type(VEC)));
  V x;
#pragma unroll
  for (int i = 0; i < VEC; ++i) x[i] = s[i];
  __builtin_nontemporal_store(x, reinterpret_cast<V*>(dst));
}

// Broadcast of lane J of each 16-lane row (ds_swizzle, bit mode: and 0x10, or J).
template <int J>
__device__ __forceinline__ int row_get(int v) {
  return __builtin_amdgcn_ds_swizzle(v, (J << 5) | 0x10);
}
template <int J>
__device__ __forceinline__ float row_get(float v) {
  return __int_as_float(row_get<J>(__float_as_int(v)));
}
template <int J>
__device__ __forceinline__ double row_get(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = row_get<J>(static_cast<int>(b & 0xffffffffll));
  const int hi = row_get<J>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// kIso: the reference's isotropic forward (csrc/forward.cu:5-53): Euclidean keys,
// w = exp(-(d^2 * s_n)) with one scale per point, w_out = sum of the K weights.
// kBwd2F: d grad_f_out of the full second order (upstream also on dp, dsigma, dR):
// sum_k w Phi f[n_k]; dq then comes from the point kernels' per-entry shares.
enum { kFwd = 0, kBwd1 = 1, kBwd2 = 2, kIso = 3, kBwd2F = 4 };

// Unrolled loop over the 16 slots of a batch (J must be a compile-time constant for
// the swizzle pattern).
template <int J, int N, typename F>
__device__ __forceinline__ void static_for(F&& fn) {
  if constexpr (J < N) {
    fn(std::integral_constant<int, J>{});
    static_for<J + 1, N>(fn);
  }
}

template <typename T, int VEC, int CH, int KEY, int MODE>
__global__ __launch_bounds__(256) void query_kernel(QueryLaunch L) {
  constexpr bool kBack = MODE == kBwd1 || MODE == kBwd2;
  const T* __restrict__ q = static_cast<const T*>(L.q);
  const T* __restrict__ p = static_cast<const T*>(L.p);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const T* __restrict__ sg = static_cast<const T*>(L.sigma);
  const T* __restrict__ R = static_cast<const T*>(L.R);
  const int64_t* __restrict__ idxs = L.idxs;
  const int64_t F = L.F;
  const int K = static_cast<int>(L.K);

  const int r = threadIdx.x & 15;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t gi = blk * 16 + (threadIdx.x >> 4);
  const bool row_ok = gi < L.rows;
  const int64_t gc = row_ok ? gi : L.rows - 1;  // keep the row's lanes active
  // spatial processing order (when given) keeps the gathered rows of neighbouring
  // queries in the same L2; results do not depend on it
  const int64_t row = L.order ? static_cast<int64_t>(L.order[gc]) : gc;
  const T* __restrict__ rec = static_cast<const T*>(L.rec);
  const int64_t b = L.cl.cloud_of_query(row);
  const int64_t pb = L.cl.p_begin(b);

  const T qx = q[row * 3 + 0], qy = q[row * 3 + 1], qz = q[row * 3 + 2];

  // feature slice of this lane: chunk c covers [(c*16 + r)*VEC, +VEC)
  T acc[CH][VEC];
  T gsl[CH][VEC];
  int64_t fofs[CH];  // clamped into the row for lanes past F
  bool fok[CH];
  const char* fbase[CH];  // this lane's slice in feature row 0 (a pointer: stays a global load)
  const uint64_t frow_bytes = static_cast<uint32_t>(F * static_cast<int64_t>(sizeof(T)));
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
    fok[c] = ft < F;
    fofs[c] = fok[c] ? ft : 0;
    fbase[c] = reinterpret_cast<const char*>(f + fofs[c]);
  }
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc[c][i] = T(0);
  if constexpr (kBack) {
    const T* __restrict__ g = static_cast<const T*>(L.grad_f_out);
    const int64_t grow = L.g_bcast ? b : row;  // broadcast upstream: one row per cloud
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
      if (ft < F) vload<T, VEC>(gsl[c], g + grow * F + ft);
      else
#pragma unroll
        for (int i = 0; i < VEC; ++i) gsl[c][i] = T(0);
    }
  }
  T hx = T(0), hy = T(0), hz = T(0);
  if constexpr (MODE == kBwd2 || MODE == kBwd2F) {
    const T* __restrict__ gg = static_cast<const T*>(L.grad_grad_q);
    if (MODE == kBwd2 || gg) {  // always given outside the full second order
      hx = gg[row * 3 + 0];
      hy = gg[row * 3 + 1];
      hz = gg[row * 3 + 2];
    }
  }
  T dqx = T(0), dqy = T(0), dqz = T(0);
  const bool want_dq = L.dq != nullptr;  // uniform: otherwise only d grad_f_out is made
  T wsum = T(0);                         // kIso: weights summed in slot order

  for (int kb = 0; kb < K; kb += 16) {
    const int k = kb + r;
    const bool have = k < K;
    const int n = have ? static_cast<int>(__builtin_nontemporal_load(idxs + row * K + k)) : 0;  // read once
    T w = T(0), coef = T(0);
    Geo<T> geo{};
    Pt<T> pt{};
    if constexpr (MODE == kIso) {
      if (have) {
        T D;
        if (L.dists) {  // the Euclidean scan's keys (same expression, same bits)
          D = static_cast<const T*>(L.dists)[row * K + k];
        } else {
          const T* __restrict__ pn = p + (pb + n) * 3;
          D = pair_key<T, AKNN_KEY_EUCLIDEAN>(qx, qy, qz, pn[0], pn[1], pn[2], T(0), T(0), T(0), T(0), T(0),
                                              T(0), T(0), T(0), T(0), T(0), T(0), T(0));
        }
        if (row_ok && L.dists_out) static_cast<T*>(L.dists_out)[row * K + k] = D;
        w = gexp(-(D * sg[pb + n]));
      }
      coef = w;
    } else if constexpr (MODE == kFwd) {
      if (have) {
        T D;
        if (L.dists && KEY == AKNN_KEY_ANISO) {  // the scan's keys are the weights' D
          D = static_cast<const T*>(L.dists)[row * K + k];
        } else {  // given neighbours or Euclidean ranking: anisotropic D, same bits as a scan
          pt = fetch_point<T>(rec, p, sg, R, pb + n);
          D = pair_key<T, AKNN_KEY_ANISO>(qx, qy, qz, pt.p0, pt.p1, pt.p2, pt.r[0], pt.r[1], pt.r[2],
                                          pt.r[3], pt.r[4], pt.r[5], pt.r[6], pt.r[7], pt.r[8], pt.s0,
                                          pt.s1, pt.s2);
          if (row_ok && L.dists_out) __builtin_nontemporal_store(D, static_cast<T*>(L.dists_out) + row * K + k);
        }
        w = gexp(-D);
        if (row_ok && L.w_out) __builtin_nontemporal_store(w, static_cast<T*>(L.w_out) + row * K + k);
      }
      coef = w;
    } else if constexpr (MODE == kBwd2F) {
      pt = fetch_point<T>(rec, p, sg, R, pb + n);
      geo = geometry<T>(qx, qy, qz, pt);
      w = have ? geo.w : T(0);
      const PtUp<T> up = load_up<T>(static_cast<const T*>(L.gg_p), static_cast<const T*>(L.gg_sigma),
                                    static_cast<const T*>(L.gg_R), pb + n);
      coef = w * second_phi<T>(geo, pt, hx, hy, hz, up).phi;  // d grad_f_out = sum_k w Phi f[n]
    } else {
      pt = fetch_point<T>(rec, p, sg, R, pb + n);
      geo = geometry<T>(qx, qy, qz, pt);
      w = have ? geo.w : T(0);
      if constexpr (MODE == kBwd2) {
        const T s = hx * geo.v0 + hy * geo.v1 + hz * geo.v2;
        coef = T(-2) * w * s;  // d grad_f_out = sum_k -2 w s f[n]
      }
    }

    if constexpr (MODE == kBwd2 || MODE == kBwd2F) {
      // d grad_w_out of slot k: the weight's upstream enters like grad_f_out . f[n_k]
      if (L.dgw && have && row_ok) static_cast<T*>(L.dgw)[row * K + k] = coef;
    }

    // All 16 row gathers of the batch are issued back to back: no per-lane branch in
    // this loop (lanes past F read a clamped in-row address and mask the value).
    int nj[16];
    T cj[16];
    static_for<0, 16>([&](auto J) {
      nj[J.value] = row_get<J.value>(n);
      cj[J.value] = row_get<J.value>(coef);
    });
    VecT<T, VEC> fv[16][CH];
    static_for<0, 16>([&](auto J) {
      // point rows < 2^31 and feature rows < 4 GB (checked by the C ABI): one 32 x 32 -> 64-bit
      // multiply-add per row address
      const uint64_t prow = static_cast<uint32_t>(pb + nj[J.value]);
#pragma unroll
      for (int c = 0; c < CH; ++c)
        fv[J.value][c] = *reinterpret_cast<const VecT<T, VEC>*>(fbase[c] + prow * frow_bytes);
    });
    if constexpr (MODE == kIso) static_for<0, 16>([&](auto J) { wsum += cj[J.value]; });
    T part[16];
    static_for<0, 16>([&](auto J) {
      T dot = T(0);
#pragma unroll
      for (int c = 0; c < CH; ++c) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) {
          // lanes past F read a clamped in-row element: the forward modes never store
          // their accumulators, the backward modes mask it out of the dot product
          const T x = kBack ? (fok[c] ? fv[J.value][c].v[i] : T(0)) : fv[J.value][c].v[i];
          if constexpr (MODE != kBwd1) acc[c][i] += cj[J.value] * x;
          if constexpr (kBack) dot += gsl[c][i] * x;
        }
      }
      part[J.value] = dot;
    });

    if (kBack && want_dq) {
      T a = row_transpose_sum<T>(part, r);  // a for slot k
      if (L.gw && have) a += static_cast<const T*>(L.gw)[row * K + k];  // grad_w_out path
      if (have) {
        if constexpr (MODE == kBwd1) {
          const T cw = T(-2) * a * w;  // dq = -2 sum a w v
          dqx += cw * geo.v0;
          dqy += cw * geo.v1;
          dqz += cw * geo.v2;
        } else {
          // dq += 2 a w (2 s v - M h), M = R^T diag(s) R
          const T s = hx * geo.v0 + hy * geo.v1 + hz * geo.v2;
          const T rh0 = pt.r[0] * hx + pt.r[1] * hy + pt.r[2] * hz;
          const T rh1 = pt.r[3] * hx + pt.r[4] * hy + pt.r[5] * hz;
          const T rh2 = pt.r[6] * hx + pt.r[7] * hy + pt.r[8] * hz;
          const T m0 = pt.s0 * rh0, m1 = pt.s1 * rh1, m2 = pt.s2 * rh2;
          const T mh0 = pt.r[0] * m0 + pt.r[3] * m1 + pt.r[6] * m2;
          const T mh1 = pt.r[1] * m0 + pt.r[4] * m1 + pt.r[7] * m2;
          const T mh2 = pt.r[2] * m0 + pt.r[5] * m1 + pt.r[8] * m2;
          const T aw2 = T(2) * a * w;
          dqx += aw2 * (T(2) * s * geo.v0 - mh0);
          dqy += aw2 * (T(2) * s * geo.v1 - mh1);
          dqz += aw2 * (T(2) * s * geo.v2 - mh2);
        }
      }
    }
  }

  if (kBack && want_dq) {
    dqx = row_sum(dqx);
    dqy = row_sum(dqy);
    dqz = row_sum(dqz);
    if (row_ok && L.dq && r < 3) {
      const T val = r == 0 ? dqx : (r == 1 ? dqy : dqz);
      static_cast<T*>(L.dq)[row * 3 + r] = val;
    }
  }
  if constexpr (MODE == kIso) {
    if (row_ok && r == 0) static_cast<T*>(L.w_out)[row] = wsum;
  }
  if constexpr (MODE != kBwd1) {
    T* __restrict__ out = static_cast<T*>((MODE == kBwd2 || MODE == kBwd2F) ? L.d_grad_f_out : L.f_out);
    if (row_ok && out) {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
        if (ft < F) {
          if (MODE == kFwd && L.accumulate) {  // add into the existing row
            T cur[VEC];
            vload<T, VEC>(cur, out + row * F + ft);
#pragma unroll
            for (int i = 0; i < VEC; ++i) cur[i] += acc[c][i];
            vstore<T, VEC>(out + row * F + ft, cur);
          } else {
            vstore_nt<T, VEC>(out + row * F + ft, acc[c]);
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ dispatch
// VEC: widest vector (16 B) that divides F and keeps every row aligned.
// CH: chunks per lane so that 16 * VEC * CH >= F.
template <typename T>
static int pick_vec(const QueryLaunch& L) {
  const int maxv = 16 / static_cast<int>(sizeof(T));
  auto aligned = [&](const void* ptr, int v) {
    return ptr == nullptr || (reinterpret_cast<uintptr_t>(ptr) % (sizeof(T) * v)) == 0;
  };
  // no wider than one chunk needs: 16 lanes x v >= F with v as small as that allows
  int v0 = 1;
  while (v0 < maxv && 16 * v0 < L.F) v0 <<= 1;
  for (int v = v0; v > 1; v >>= 1) {
    if (L.F % v == 0 && aligned(L.f, v) && aligned(L.f_out, v) && aligned(L.grad_f_out, v) &&
        aligned(L.d_grad_f_out, v))
      return v;
  }
  return 1;
}

template <typename T, int VEC, int KEY, int MODE>
static hipError_t launch_q_ch(const QueryLaunch& L, hipStream_t st) {
  const int64_t per = 16 * VEC;
  const int64_t ch = (L.F + per - 1) / per;
  const dim3 grid(static_cast<unsigned>((L.rows + 15) / 16)), block(256);
  if (ch <= 1) hipLaunchKernelGGL((query_kernel<T, VEC, 1, KEY, MODE>), grid, block, 0, st, L);
  else if (ch <= 2) hipLaunchKernelGGL((query_kernel<T, VEC, 2, KEY, MODE>), grid, block, 0, st, L);
  else if (ch <= 4) hipLaunchKernelGGL((query_kernel<T, VEC, 4, KEY, MODE>), grid, block, 0, st, L);
  else if (ch <= 8) hipLaunchKernelGGL((query_kernel<T, VEC, 8, KEY, MODE>), grid, block, 0, st, L);
  else return hipErrorNotSupported;  // F wider than 16 * VEC * 8
  return hipGetLastError();
}

template <typename T, int KEY, int MODE>
static hipError_t launch_q_vec(const QueryLaunch& L, hipStream_t st) {
  const int v = pick_vec<T>(L);
  if constexpr (sizeof(T) == 4) {
    if (v == 4) return launch_q_ch<T, 4, KEY, MODE>(L, st);
  }
  if (v == 2) return launch_q_ch<T, 2, KEY, MODE>(L, st);
  return launch_q_ch<T, 1, KEY, MODE>(L, st);
}

template <int MODE>
static hipError_t launch_q(const QueryLaunch& L, hipStream_t st) {
  if (L.rows == 0) return hipSuccess;
  if (L.dtype == AKNN_F32)
    return L.key == AKNN_KEY_EUCLIDEAN ? launch_q_vec<float, AKNN_KEY_EUCLIDEAN, MODE>(L, st)
                                       : launch_q_vec<float, AKNN_KEY_ANISO, MODE>(L, st);
  return L.key == AKNN_KEY_EUCLIDEAN ? launch_q_vec<double, AKNN_KEY_EUCLIDEAN, MODE>(L, st)
                                     : launch_q_vec<double, AKNN_KEY_ANISO, MODE>(L, st);
}

hipError_t launch_aggregate(const QueryLaunch& L, hipStream_t st) { return launch_q<kFwd>(L, st); }

hipError_t launch_aggregate_iso(const QueryLaunch& L, hipStream_t st) {
  if (L.dtype == AKNN_F32) return launch_q_vec<float, AKNN_KEY_EUCLIDEAN, kIso>(L, st);
  return launch_q_vec<double, AKNN_KEY_EUCLIDEAN, kIso>(L, st);
}

// ------------------------------------------------------------------ lookup
// The reference's table lookup (csrc/forward.cu:113-183): the query's Euclidean
// nearest point n1 (first index reaching the minimum below 1e10, else 0) selects row
// n1 of the [P, table_k] neighbour table; idxs[q, k] = table[n1, k] for k < K.
template <typename T>
__global__ void lookup_gather_kernel(const T* __restrict__ d1, const int64_t* __restrict__ i1,
                                     const int64_t* __restrict__ table, int64_t table_k, int64_t rows,
                                     int K, int64_t* __restrict__ idxs) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * K) return;
  const int64_t row = i / K;
  const int k = static_cast<int>(i - row * K);
  const int64_t n1 = d1[row] < static_cast<T>(1e10) ? i1[row] : 0;
  idxs[i] = table[n1 * table_k + k];
}

hipError_t launch_lookup_gather(const void* d1, const int64_t* i1, const int64_t* table, int64_t table_k,
                                int64_t rows, int64_t K, int32_t dtype, int64_t* idxs, hipStream_t st) {
  const int64_t n = rows * K;
  if (n == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
  if (dtype == AKNN_F32)
    hipLaunchKernelGGL(lookup_gather_kernel<float>, grid, block, 0, st, static_cast<const float*>(d1), i1, table,
                       table_k, rows, static_cast<int>(K), idxs);
  else
    hipLaunchKernelGGL(lookup_gather_kernel<double>, grid, block, 0, st, static_cast<const double*>(d1), i1,
                       table, table_k, rows, static_cast<int>(K), idxs);
  return hipGetLastError();
}

hipError_t launch_query_grad_full(const QueryLaunch& L, hipStream_t st) { return launch_q<kBwd2F>(L, st); }

// ------------------------------------------------------------------ dq from point dots
// Broadcast upstream: a = g . f[n] depends on the point alone and comes precomputed in
// the dot records (L.dots: {p, R, s, a} per point row), so dq needs no feature rows at
// all -- per query only its idxs row, the K dot records (L2-resident) and, in the second
// order, grad_grad_q.  One 16-lane row per query, lane r on slot kb + r.  Each slot's
// share is the point kernels' per-entry share expression; the 16 shares of a batch are
// summed by a fixed DPP tree over the row (deterministic), batches in slot order.
//
// A lane-per-record gather would touch a different cache line in every lane of every
// load and keep the texture addresser busy for the whole kernel.  Instead each wave
// loads its 64 slots' records cooperatively -- 16-byte chunks, the lanes of one record
// adjacent, so each load covers 64 / CPR whole records -- into LDS, and each lane then
// reads its own record from there.
template <typename T>
__device__ __forceinline__ void wave_records(const T* __restrict__ rec, int gp, T (*stage)[16], Pt<T>& o,
                                             T& a) {
  constexpr int kCpr = static_cast<int>(sizeof(T));  // 16-byte chunks per record (16 values)
  constexpr int kVals = 16 / kCpr;                     // values per chunk
  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;
  __syncthreads();  // the previous batch's reads of this block's stage are done
#pragma unroll
  for (int j = 0; j < kCpr; ++j) {
    const int s = j * (64 / kCpr) + lane / kCpr;  // slot (= lane of the wave) whose record this lane fetches
    const int c = lane % kCpr;
    const int gs = __shfl(gp, s, 64);
    const T* src = rec + static_cast<int64_t>(gs) * 16 + c * kVals;
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<float4*>(&stage[wbase + s][c * kVals]) = *reinterpret_cast<const float4*>(src);
    } else {
      *reinterpret_cast<double2*>(&stage[wbase + s][c * kVals]) = *reinterpret_cast<const double2*>(src);
    }
  }
  __syncthreads();
  const T* v = stage[threadIdx.x];
  o.p0 = v[0];
  o.p1 = v[1];
  o.p2 = v[2];
#pragma unroll
  for (int i = 0; i < 9; ++i) o.r[i] = v[3 + i];
  o.s0 = v[12];
  o.s1 = v[13];
  o.s2 = v[14];
  a = v[15];
}

template <typename T, bool SECOND>
__global__ __launch_bounds__(256) void dq_dots_kernel(QueryLaunch L) {
  __shared__ __align__(16) T stage[256][16];  // one dot record per slot of the block's four waves
  const T* __restrict__ q = static_cast<const T*>(L.q);
  const T* __restrict__ dots = static_cast<const T*>(L.dots);
  const int64_t* __restrict__ idxs = L.idxs;
  const int K = static_cast<int>(L.K);

  const int r = threadIdx.x & 15;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t gi = blk * 16 + (threadIdx.x >> 4);
  const bool row_ok = gi < L.rows;
  const int64_t gc = row_ok ? gi : L.rows - 1;
  const int64_t row = L.order ? static_cast<int64_t>(L.order[gc]) : gc;
  const int64_t pb = L.cl.p_begin(L.cl.cloud_of_query(row));
  const T qx = q[row * 3 + 0], qy = q[row * 3 + 1], qz = q[row * 3 + 2];
  T hx = T(0), hy = T(0), hz = T(0);
  if constexpr (SECOND) {
    const T* __restrict__ hh = static_cast<const T*>(L.grad_grad_q);
    hx = hh[row * 3 + 0];
    hy = hh[row * 3 + 1];
    hz = hh[row * 3 + 2];
  }
  T s0 = T(0), s1 = T(0), s2 = T(0);
  for (int kb = 0; kb < K; kb += 16) {
    const int k = kb + r;
    const bool have = k < K;
    const int n = have ? static_cast<int>(__builtin_nontemporal_load(idxs + row * K + k)) : 0;
    const int gp = static_cast<int>(pb) + n;  // point rows < 2^31 (checked by the C ABI)
    Pt<T> pt;
    T a;
    wave_records<T>(dots, gp, stage, pt, a);
    if (L.gw && have) a += static_cast<const T*>(L.gw)[row * K + k];  // grad_w_out path
    const Geo<T> geo = geometry<T>(qx, qy, qz, pt);
    const T w = have ? geo.w : T(0);
    const T aw = a * w;
    T c0, c1, c2;
    if constexpr (!SECOND) {
      c0 = T(-2) * aw * geo.v0;
      c1 = T(-2) * aw * geo.v1;
      c2 = T(-2) * aw * geo.v2;
    } else {
      const T s = hx * geo.v0 + hy * geo.v1 + hz * geo.v2;
      if (L.dgw && have && row_ok) static_cast<T*>(L.dgw)[row * K + k] = T(-2) * w * s;  // d grad_w_out
      T hr[3];  // h . R_i
#pragma unroll
      for (int i = 0; i < 3; ++i) hr[i] = pt.r[3 * i] * hx + pt.r[3 * i + 1] * hy + pt.r[3 * i + 2] * hz;
      const T m0 = pt.s0 * hr[0], m1 = pt.s1 * hr[1], m2 = pt.s2 * hr[2];
      const T mh0 = pt.r[0] * m0 + pt.r[3] * m1 + pt.r[6] * m2;
      const T mh1 = pt.r[1] * m0 + pt.r[4] * m1 + pt.r[7] * m2;
      const T mh2 = pt.r[2] * m0 + pt.r[5] * m1 + pt.r[8] * m2;
      const T t0 = T(2) * s * geo.v0 - mh0, t1 = T(2) * s * geo.v1 - mh1, t2 = T(2) * s * geo.v2 - mh2;
      c0 = T(2) * aw * t0;
      c1 = T(2) * aw * t1;
      c2 = T(2) * aw * t2;
    }
    // the batch's slot shares summed over the row by a fixed DPP tree (slots past K add
    // zero), batches in order
    s0 += row_sum(have ? c0 : T(0));
    s1 += row_sum(have ? c1 : T(0));
    s2 += row_sum(have ? c2 : T(0));
  }
  if (row_ok && r < 3) static_cast<T*>(L.dq)[row * 3 + r] = r == 0 ? s0 : (r == 1 ? s1 : s2);
}

template <typename T, bool SECOND>
static void launch_dq_dots_t(const QueryLaunch& L, dim3 grid, dim3 block, hipStream_t st) {
  hipLaunchKernelGGL((dq_dots_kernel<T, SECOND>), grid, block, 0, st, L);
}

hipError_t launch_dq_dots(const QueryLaunch& L, bool second_order, hipStream_t st) {
  if (L.rows == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((L.rows + 15) / 16)), block(256);
  if (L.dtype == AKNN_F32) {
    if (second_order) launch_dq_dots_t<float, true>(L, grid, block, st);
    else launch_dq_dots_t<float, false>(L, grid, block, st);
  } else {
    if (second_order) launch_dq_dots_t<double, true>(L, grid, block, st);
    else launch_dq_dots_t<double, false>(L, grid, block, st);
  }
  return hipGetLastError();
}

hipError_t launch_query_grad(const QueryLaunch& L, bool second_order, hipStream_t st) {
  return second_order ? launch_q<kBwd2>(L, st) : launch_q<kBwd1>(L, st);
}

}  // namespace aknn
