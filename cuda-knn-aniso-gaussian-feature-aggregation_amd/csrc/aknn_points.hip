// Point-side gradients for gfx950 without float atomics.
//
// The reference scatters every (query, slot) contribution into the point arrays with
// float atomicAdd (csrc/backward.cu:91,102-119 first order; :291,308-326 second
// order), so its sums depend on thread timing.  Here the neighbour table is
// transposed once into a CSR over point rows (entries q*K+k in ascending order, built
// by a stable radix sort), and one 16-lane row per point walks its entries in that
// fixed order: every sum has one fixed association, so results are bitwise
// reproducible.
//
// Per entry (query q, slot k, point n), recomputed from idxs (never stored):
//   a = grad_f_out[q] . f[n],   geometry u, d, w, v as in the forward
// First order  (csrc/backward.cu:64-119):
//   dp += 2 a w v;  dsigma_i += -a w u_i^2;  dR_ij += -2 a w s_i u_i d_j;  df += w g
// Second order with h = grad_grad_q[q], s = h.v, t = 2 s v - M h (csrc/backward.cu:196-326):
//   dp += -2 a w t;  dsigma_i += 2 a w u_i (u_i s - h.R_i);
//   dR_ij += -2 a w s_i (h_j u_i + (h.R_i) d_j - 2 s u_i d_j);  df += -2 w s g
#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

template <typename T, int VEC>
struct alignas(sizeof(T) * VEC) PVec {
  T v[VEC];
};

template <int J>
__device__ __forceinline__ int prow_get(int v) {
  return __builtin_amdgcn_ds_swizzle(v, (J << 5) | 0x10);
}
template <int J>
__device__ __forceinline__ float prow_get(float v) {
  return __int_as_float(prow_get<J>(__float_as_int(v)));
}
template <int J>
__device__ __forceinline__ double prow_get(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = prow_get<J>(static_cast<int>(b & 0xffffffffll));
  const int hi = prow_get<J>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

template <int J, int N, typename Fn>
__device__ __forceinline__ void pstatic_for(Fn&& fn) {
  if constexpr (J < N) {
    fn(std::integral_constant<int, J>{});
    pstatic_for<J + 1, N>(fn);
  }
}

// With one point row per wave (G = 4) the row index is the same in every lane: said so
// to the compiler, the row's CSR bounds, point data and cloud lookup become scalar
// loads and the entry loop a scalar loop.
template <bool UNIFORM>
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
  if constexpr (UNIFORM) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return v;
  }
}

// Sum over the G 16-lane groups of a wave that share one point row (fixed order:
// (g0 + g1) + (g2 + g3); every group ends with the same bits).
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  if constexpr (G >= 2) v = v + __shfl_xor(v, 16, 64);
  if constexpr (G >= 4) v = v + __shfl_xor(v, 32, 64);
  return v;
}

// Broadcast upstream (one grad_f_out row per cloud, BC): this lane's slice of the
// point row's cloud row, and a = g . f[row] reduced across the row in the association
// of the per-entry dot products (row_transpose_sum over 16 equal partials: the same
// pairing tree as for any one entry), so a, and with it dp, dsigma and dR, keep the
// bits of a materialised upstream.
template <typename T, int VEC, int CH>
__device__ __forceinline__ T bcast_row_dot(const T* __restrict__ g, int64_t b, int64_t F, const int64_t (&fofs)[CH],
                                           const bool (&fok)[CH], const T (&fsl)[CH][VEC], T (&gbc)[CH][VEC],
                                           int r) {
  T dot = T(0);
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const PVec<T, VEC> x = *reinterpret_cast<const PVec<T, VEC>*>(g + b * F + fofs[c]);
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      gbc[c][i] = fok[c] ? x.v[i] : T(0);
      dot += fsl[c][i] * gbc[c][i];
    }
  }
  T part[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) part[j] = dot;
  return row_transpose_sum<T>(part, r);
}

// G = 16-lane groups per point row (1: four rows per wave; 4: one row per wave, its
// entries dealt to the groups in turn).  Fewer rows in flight keep the gathered query
// rows of the rows being processed within L2.  BC: broadcast upstream (no row gathers).
template <typename T, int VEC, int CH, int KEY, bool SECOND, int G, bool BC>
__global__ __launch_bounds__(256) void point_kernel(PointLaunch L) {
  const T* __restrict__ q = static_cast<const T*>(L.q);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const T* __restrict__ g = static_cast<const T*>(L.grad_f_out);
  const T* __restrict__ hh = static_cast<const T*>(L.grad_grad_q);
  const int32_t* __restrict__ off = L.csr_off;
  const int32_t* __restrict__ ent = L.csr_ent;
  const int64_t F = L.F;
  const int K = static_cast<int>(L.K);
  const FastDiv kdiv(static_cast<uint32_t>(K));  // entry -> query row (entries < 2^31)
  // CSR in the spatial query order: query-side inputs from the ordered copies, and the
  // query row itself only where a per-query array is indexed by it
  const bool ordered = L.qorder != nullptr;
  const T* __restrict__ qsrc = ordered ? static_cast<const T*>(L.q_sorted) : q;
  const T* __restrict__ hsrc = ordered ? static_cast<const T*>(L.hh_sorted) : hh;
  const bool need_row = !BC || L.dq_contrib != nullptr || L.gw != nullptr;  // uniform

  const int r = threadIdx.x & 15;
  const int gq = (threadIdx.x >> 4) % G;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t gi = wave_uniform<G == 4>(blk * (16 / G) + (threadIdx.x >> 4) / G);
  const bool row_ok = gi < L.rows;
  const int64_t gc = row_ok ? gi : L.rows - 1;
  // spatial processing order (when given): neighbouring points share queries, so the
  // gathered grad_f_out rows stay in L2; results do not depend on it
  const int64_t row = L.order ? static_cast<int64_t>(L.order[gc]) : gc;

  const Pt<T> pt = load_point<T>(static_cast<const T*>(L.p), static_cast<const T*>(L.sigma),
                                 static_cast<const T*>(L.R), row);
  const int e0 = off[row];
  const int e1 = row_ok ? off[row + 1] : e0;  // same trip count for the whole row

  T fsl[CH][VEC], dfa[CH][VEC];
  int64_t fofs[CH];  // this lane's feature offset per chunk, clamped into the row
  bool fok[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
    fok[c] = ft < F;
    fofs[c] = fok[c] ? ft : 0;
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      dfa[c][i] = T(0);
      fsl[c][i] = T(0);
    }
    if (ft < F) {
      const PVec<T, VEC> x = *reinterpret_cast<const PVec<T, VEC>*>(f + row * F + ft);
#pragma unroll
      for (int i = 0; i < VEC; ++i) fsl[c][i] = x.v[i];
    }
  }
  T acc[15];  // dp[3], dsigma[3], dR[9]
#pragma unroll
  for (int i = 0; i < 15; ++i) acc[i] = T(0);
  T gbc[BC ? CH : 1][VEC];
  T a_row = T(0);
  if constexpr (BC) a_row = bcast_row_dot<T, VEC, CH>(g, L.cl.cloud_of_point(row), F, fofs, fok, fsl, gbc, r);
  // broadcast upstream: every entry's df term is coef * (the cloud's row), so df is the
  // row times the sum of the coefficients (each lane sums its own entries, in order)
  T csum = T(0);

  for (int eb = e0; eb < e1; eb += 16 * G) {
    const int e = eb + 16 * gq + r;
    const bool have = e < e1;
    const int entry = have ? ent[e] : 0;
    // the entry's query: position pos of the CSR's numbering, row qrow, slot entry
    // qrow*K + k (pos = qrow unless the CSR is in the spatial query order)
    const int64_t pos = have ? static_cast<int64_t>(kdiv.div(static_cast<uint32_t>(entry))) : 0;
    const int64_t qrow = (ordered && need_row) ? static_cast<int64_t>(L.qorder[pos]) : pos;
    const int oentry = ordered ? static_cast<int>(qrow * K + (entry - pos * K)) : entry;
    const T qx = qsrc[pos * 3 + 0], qy = qsrc[pos * 3 + 1], qz = qsrc[pos * 3 + 2];
    const Geo<T> geo = geometry<T>(qx, qy, qz, pt);
    const T w = have ? geo.w : T(0);
    T hx = T(0), hy = T(0), hz = T(0), s = T(0), coef;
    if constexpr (SECOND) {
      hx = hsrc[pos * 3 + 0];
      hy = hsrc[pos * 3 + 1];
      hz = hsrc[pos * 3 + 2];
      s = hx * geo.v0 + hy * geo.v1 + hz * geo.v2;
      coef = T(-2) * w * s;  // df += -2 w s g
    } else {
      coef = w;  // df += w g
    }
    const int qi32 = static_cast<int>(qrow);

    T a;
    if constexpr (BC) {
      csum += coef;  // zero for lanes without an entry (w = 0)
      a = a_row;
      (void)qi32;
    } else {
      T cj[16];
      pstatic_for<0, 16>([&](auto J) { cj[J.value] = prow_get<J.value>(coef); });
      // All 16 row gathers of the batch are issued back to back: no per-lane branch in
      // this loop (lanes past F read a clamped in-row address and mask the value).
      int qj[16];
      pstatic_for<0, 16>([&](auto J) { qj[J.value] = prow_get<J.value>(qi32); });
      PVec<T, VEC> xs[16][CH];
      pstatic_for<0, 16>([&](auto J) {
        const T* __restrict__ grow = g + static_cast<int64_t>(qj[J.value]) * F;
#pragma unroll
        for (int c = 0; c < CH; ++c) xs[J.value][c] = *reinterpret_cast<const PVec<T, VEC>*>(grow + fofs[c]);
      });
      T part[16];
      pstatic_for<0, 16>([&](auto J) {
        T dot = T(0);
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
          for (int i = 0; i < VEC; ++i) {
            const T x = fok[c] ? xs[J.value][c].v[i] : T(0);
            dfa[c][i] += cj[J.value] * x;
            dot += fsl[c][i] * x;
          }
        }
        part[J.value] = dot;
      });
      a = row_transpose_sum<T>(part, r);
    }
    if (L.gw && have) a += static_cast<const T*>(L.gw)[oentry];  // grad_w_out path

    if (have) {
      const T u[3] = {geo.u0, geo.u1, geo.u2};
      const T d[3] = {geo.d0, geo.d1, geo.d2};
      const T sg[3] = {pt.s0, pt.s1, pt.s2};
      const T aw = a * w;
      T* __restrict__ cq = static_cast<T*>(L.dq_contrib);
      const int64_t cslot = static_cast<int64_t>(oentry);
      if constexpr (!SECOND) {
        acc[0] += T(2) * aw * geo.v0;
        acc[1] += T(2) * aw * geo.v1;
        acc[2] += T(2) * aw * geo.v2;
        if (cq) {  // this entry's share of dq (= minus its share of dp)
          cq[cslot * 3 + 0] = T(-2) * aw * geo.v0;
          cq[cslot * 3 + 1] = T(-2) * aw * geo.v1;
          cq[cslot * 3 + 2] = T(-2) * aw * geo.v2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          acc[3 + i] += -aw * u[i] * u[i];
          const T cr = T(-2) * aw * sg[i] * u[i];
#pragma unroll
          for (int j = 0; j < 3; ++j) acc[6 + 3 * i + j] += cr * d[j];
        }
      } else {
        const T h[3] = {hx, hy, hz};
        T hr[3];  // h . R_i
#pragma unroll
        for (int i = 0; i < 3; ++i) hr[i] = pt.r[3 * i] * hx + pt.r[3 * i + 1] * hy + pt.r[3 * i + 2] * hz;
        const T m0 = pt.s0 * hr[0], m1 = pt.s1 * hr[1], m2 = pt.s2 * hr[2];
        const T mh0 = pt.r[0] * m0 + pt.r[3] * m1 + pt.r[6] * m2;
        const T mh1 = pt.r[1] * m0 + pt.r[4] * m1 + pt.r[7] * m2;
        const T mh2 = pt.r[2] * m0 + pt.r[5] * m1 + pt.r[8] * m2;
        const T aw2 = T(-2) * aw;
        const T t0 = T(2) * s * geo.v0 - mh0, t1 = T(2) * s * geo.v1 - mh1, t2 = T(2) * s * geo.v2 - mh2;
        acc[0] += aw2 * t0;
        acc[1] += aw2 * t1;
        acc[2] += aw2 * t2;
        if (cq) {  // this entry's share of dq: 2 a w (2 s v - M h)
          cq[cslot * 3 + 0] = T(2) * aw * t0;
          cq[cslot * 3 + 1] = T(2) * aw * t1;
          cq[cslot * 3 + 2] = T(2) * aw * t2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          acc[3 + i] += T(2) * aw * u[i] * (u[i] * s - hr[i]);
          const T cr = aw2 * sg[i];
#pragma unroll
          for (int j = 0; j < 3; ++j)
            acc[6 + 3 * i + j] += cr * (h[j] * u[i] + hr[i] * d[j] - T(2) * s * u[i] * d[j]);
        }
      }
    }
  }

#pragma unroll
  for (int i = 0; i < 15; ++i) acc[i] = group_sum<G>(row_sum(acc[i]));
  if constexpr (BC) {
    const T tot = group_sum<G>(row_sum(csum));
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int i = 0; i < VEC; ++i) dfa[c][i] = tot * gbc[c][i];
  } else if constexpr (G > 1) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int i = 0; i < VEC; ++i) dfa[c][i] = group_sum<G>(dfa[c][i]);
  }
  if (!row_ok || gq != 0) return;
  T* dp = static_cast<T*>(L.dp);
  T* ds = static_cast<T*>(L.dsigma);
  T* dR = static_cast<T*>(L.dR);
  T* df = static_cast<T*>(L.df);
  // lane r < 15 stores accumulator r
  T mine = acc[0];
#pragma unroll
  for (int i = 1; i < 15; ++i) mine = (r == i) ? acc[i] : mine;
  if (r < 3) {
    if (dp) dp[row * 3 + r] = mine;
  } else if (r < 6) {
    if (ds) ds[row * 3 + (r - 3)] = mine;
  } else if (r < 15) {
    if (dR) dR[row * 9 + (r - 6)] = mine;
  }
  if (df) {
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
      if (ft < F) {
        PVec<T, VEC> x;
#pragma unroll
        for (int i = 0; i < VEC; ++i) x.v[i] = dfa[c][i];
        *reinterpret_cast<PVec<T, VEC>*>(df + row * F + ft) = x;
      }
    }
  }
}

// The full second order as a kernel of its own (the reference-order kernel above stays
// as it is): the upstream also holds gradients of dp, df, dsigma, dR (L.gg_p / gg_f /
// gg_sigma / gg_R, each optional).  Instantiated with SECOND = FULL = true.
template <typename T, int VEC, int CH, int KEY, bool SECOND, int G, bool FULL>
__global__ __launch_bounds__(256) void point_kernel_full(PointLaunch L) {
  const T* __restrict__ q = static_cast<const T*>(L.q);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const T* __restrict__ g = static_cast<const T*>(L.grad_f_out);
  const T* __restrict__ hh = static_cast<const T*>(L.grad_grad_q);  // may be null with FULL
  const int32_t* __restrict__ off = L.csr_off;
  const int32_t* __restrict__ ent = L.csr_ent;
  const int64_t F = L.F;
  const int K = static_cast<int>(L.K);
  const FastDiv kdiv(static_cast<uint32_t>(K));  // entry -> query row (entries < 2^31)
  // CSR in the spatial query order: query-side inputs from the ordered copies, and the
  // query row itself only where a per-query array is indexed by it
  const bool ordered = L.qorder != nullptr;
  const T* __restrict__ qsrc = ordered ? static_cast<const T*>(L.q_sorted) : q;
  const T* __restrict__ hsrc = ordered ? static_cast<const T*>(L.hh_sorted) : hh;
  const bool need_row = true;  // grad_f_out rows are gathered per entry

  const int r = threadIdx.x & 15;
  const int gq = (threadIdx.x >> 4) % G;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t gi = wave_uniform<G == 4>(blk * (16 / G) + (threadIdx.x >> 4) / G);
  const bool row_ok = gi < L.rows;
  const int64_t gc = row_ok ? gi : L.rows - 1;
  // spatial processing order (when given): neighbouring points share queries, so the
  // gathered grad_f_out rows stay in L2; results do not depend on it
  const int64_t row = L.order ? static_cast<int64_t>(L.order[gc]) : gc;

  const Pt<T> pt = load_point<T>(static_cast<const T*>(L.p), static_cast<const T*>(L.sigma),
                                 static_cast<const T*>(L.R), row);
  const int e0 = off[row];
  const int e1 = row_ok ? off[row + 1] : e0;  // same trip count for the whole row

  T fsl[CH][VEC], dfa[CH][VEC];
  int64_t fofs[CH];  // this lane's feature offset per chunk, clamped into the row
  bool fok[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
    fok[c] = ft < F;
    fofs[c] = fok[c] ? ft : 0;
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      dfa[c][i] = T(0);
      fsl[c][i] = T(0);
    }
    if (ft < F) {
      const PVec<T, VEC> x = *reinterpret_cast<const PVec<T, VEC>*>(f + row * F + ft);
#pragma unroll
      for (int i = 0; i < VEC; ++i) fsl[c][i] = x.v[i];
    }
  }
  T acc[15];  // dp[3], dsigma[3], dR[9]
#pragma unroll
  for (int i = 0; i < 15; ++i) acc[i] = T(0);
  // full second order: this point's upstream (dp, dsigma, dR parts) and gg_f slice
  PtUp<T> up{};
  T hfs[FULL ? CH : 1][VEC];
  const bool has_hf = FULL && L.gg_f != nullptr;  // uniform
  if constexpr (FULL) {
    up = load_up<T>(static_cast<const T*>(L.gg_p), static_cast<const T*>(L.gg_sigma),
                    static_cast<const T*>(L.gg_R), row);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) hfs[c][i] = T(0);
      if (has_hf && fok[c]) {
        const PVec<T, VEC> x =
            *reinterpret_cast<const PVec<T, VEC>*>(static_cast<const T*>(L.gg_f) + row * F + fofs[c]);
#pragma unroll
        for (int i = 0; i < VEC; ++i) hfs[c][i] = x.v[i];
      }
    }
  }

  for (int eb = e0; eb < e1; eb += 16 * G) {
    const int e = eb + 16 * gq + r;
    const bool have = e < e1;
    const int entry = have ? ent[e] : 0;
    // the entry's query: position pos of the CSR's numbering, row qrow, slot entry
    // qrow*K + k (pos = qrow unless the CSR is in the spatial query order)
    const int64_t pos = have ? static_cast<int64_t>(kdiv.div(static_cast<uint32_t>(entry))) : 0;
    const int64_t qrow = (ordered && need_row) ? static_cast<int64_t>(L.qorder[pos]) : pos;
    const int oentry = ordered ? static_cast<int>(qrow * K + (entry - pos * K)) : entry;
    const T qx = qsrc[pos * 3 + 0], qy = qsrc[pos * 3 + 1], qz = qsrc[pos * 3 + 2];
    const Geo<T> geo = geometry<T>(qx, qy, qz, pt);
    const T w = have ? geo.w : T(0);
    T hx = T(0), hy = T(0), hz = T(0), s = T(0), coef;
    Second2<T> sec{};
    if constexpr (SECOND) {
      // grad_grad_q is always given outside the full second order: loaded with no
      // branch there, so these loads issue together with the query's
      if (!FULL || hh) {
        hx = hsrc[pos * 3 + 0];
        hy = hsrc[pos * 3 + 1];
        hz = hsrc[pos * 3 + 2];
      }
      if constexpr (FULL) {
        sec = second_phi<T>(geo, pt, hx, hy, hz, up);
        coef = w * sec.phi;  // df += w Phi g
      } else {
        s = hx * geo.v0 + hy * geo.v1 + hz * geo.v2;
        coef = T(-2) * w * s;  // df += -2 w s g
      }
    } else {
      coef = w;  // df += w g
    }
    const int qi32 = static_cast<int>(qrow);

    // All 16 row gathers of the batch are issued back to back: no per-lane branch in
    // this loop (lanes past F read a clamped in-row address and mask the value).
    int qj[16];
    T cj[16];
    pstatic_for<0, 16>([&](auto J) {
      qj[J.value] = prow_get<J.value>(qi32);
      cj[J.value] = prow_get<J.value>(coef);
    });
    PVec<T, VEC> xs[16][CH];
    pstatic_for<0, 16>([&](auto J) {
      const T* __restrict__ grow = g + static_cast<int64_t>(qj[J.value]) * F;
#pragma unroll
      for (int c = 0; c < CH; ++c) xs[J.value][c] = *reinterpret_cast<const PVec<T, VEC>*>(grow + fofs[c]);
    });
    T part[16], part2[FULL ? 16 : 1];
    pstatic_for<0, 16>([&](auto J) {
      T dot = T(0), dot2 = T(0);
#pragma unroll
      for (int c = 0; c < CH; ++c) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) {
          const T x = fok[c] ? xs[J.value][c].v[i] : T(0);
          dfa[c][i] += cj[J.value] * x;
          dot += fsl[c][i] * x;
          if constexpr (FULL) dot2 += hfs[c][i] * x;
        }
      }
      part[J.value] = dot;
      if constexpr (FULL) part2[J.value] = dot2;
    });
    T a = row_transpose_sum<T>(part, r);
    if (L.gw && have) a += static_cast<const T*>(L.gw)[oentry];  // grad_w_out path
    T ghf = T(0);  // g[q] . gg_f[n]
    if constexpr (FULL) {
      if (has_hf) ghf = row_transpose_sum<T>(part2, r);
    }

    if (have) {
      const T u[3] = {geo.u0, geo.u1, geo.u2};
      const T d[3] = {geo.d0, geo.d1, geo.d2};
      const T sg[3] = {pt.s0, pt.s1, pt.s2};
      const T aw = a * w;
      T* __restrict__ cq = static_cast<T*>(L.dq_contrib);
      const int64_t cslot = static_cast<int64_t>(oentry);
      if constexpr (!SECOND) {
        acc[0] += T(2) * aw * geo.v0;
        acc[1] += T(2) * aw * geo.v1;
        acc[2] += T(2) * aw * geo.v2;
        if (cq) {  // this entry's share of dq (= minus its share of dp)
          cq[cslot * 3 + 0] = T(-2) * aw * geo.v0;
          cq[cslot * 3 + 1] = T(-2) * aw * geo.v1;
          cq[cslot * 3 + 2] = T(-2) * aw * geo.v2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          acc[3 + i] += -aw * u[i] * u[i];
          const T cr = T(-2) * aw * sg[i] * u[i];
#pragma unroll
          for (int j = 0; j < 3; ++j) acc[6 + 3 * i + j] += cr * d[j];
        }
      } else if constexpr (FULL) {
        // c = a Phi + g.Hf;  d c_e/dd = -2 w c v + a w dPhi/dd, with
        // dPhi/dd = -2 [R^T t + HR^T (sig*u)],  t_i = sig_i (R_i.hh) + Hs_i u_i + sig_i (HR_i.d)
        const T c = a * sec.phi + ghf;
        T t[3], su[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          t[i] = sg[i] * sec.rh[i] + up.s[i] * u[i] + sg[i] * sec.hrd[i];
          su[i] = sg[i] * u[i];
        }
        const T vv[3] = {geo.v0, geo.v1, geo.v2};
        T gd[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const T dphi = T(-2) * (pt.r[j] * t[0] + pt.r[3 + j] * t[1] + pt.r[6 + j] * t[2] +
                                  up.r[j] * su[0] + up.r[3 + j] * su[1] + up.r[6 + j] * su[2]);
          gd[j] = T(-2) * w * c * vv[j] + aw * dphi;
          acc[j] -= gd[j];  // dp = -d/dd
        }
        if (cq) {  // this entry's share of dq (= +d/dd)
          cq[cslot * 3 + 0] = gd[0];
          cq[cslot * 3 + 1] = gd[1];
          cq[cslot * 3 + 2] = gd[2];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          acc[3 + i] += -w * u[i] * (u[i] * c + T(2) * a * (sec.rh[i] + sec.hrd[i]));
#pragma unroll
          for (int j = 0; j < 3; ++j)
            acc[6 + 3 * i + j] += T(-2) * w * (su[i] * d[j] * c + a * (su[i] * sec.hh[j] + d[j] * t[i]));
        }
      } else {
        const T h[3] = {hx, hy, hz};
        T hr[3];  // h . R_i
#pragma unroll
        for (int i = 0; i < 3; ++i) hr[i] = pt.r[3 * i] * hx + pt.r[3 * i + 1] * hy + pt.r[3 * i + 2] * hz;
        const T m0 = pt.s0 * hr[0], m1 = pt.s1 * hr[1], m2 = pt.s2 * hr[2];
        const T mh0 = pt.r[0] * m0 + pt.r[3] * m1 + pt.r[6] * m2;
        const T mh1 = pt.r[1] * m0 + pt.r[4] * m1 + pt.r[7] * m2;
        const T mh2 = pt.r[2] * m0 + pt.r[5] * m1 + pt.r[8] * m2;
        const T aw2 = T(-2) * aw;
        const T t0 = T(2) * s * geo.v0 - mh0, t1 = T(2) * s * geo.v1 - mh1, t2 = T(2) * s * geo.v2 - mh2;
        acc[0] += aw2 * t0;
        acc[1] += aw2 * t1;
        acc[2] += aw2 * t2;
        if (cq) {  // this entry's share of dq: 2 a w (2 s v - M h)
          cq[cslot * 3 + 0] = T(2) * aw * t0;
          cq[cslot * 3 + 1] = T(2) * aw * t1;
          cq[cslot * 3 + 2] = T(2) * aw * t2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          acc[3 + i] += T(2) * aw * u[i] * (u[i] * s - hr[i]);
          const T cr = aw2 * sg[i];
#pragma unroll
          for (int j = 0; j < 3; ++j)
            acc[6 + 3 * i + j] += cr * (h[j] * u[i] + hr[i] * d[j] - T(2) * s * u[i] * d[j]);
        }
      }
    }
  }

#pragma unroll
  for (int i = 0; i < 15; ++i) acc[i] = group_sum<G>(row_sum(acc[i]));
  if constexpr (G > 1) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int i = 0; i < VEC; ++i) dfa[c][i] = group_sum<G>(dfa[c][i]);
  }
  if (!row_ok || gq != 0) return;
  T* dp = static_cast<T*>(L.dp);
  T* ds = static_cast<T*>(L.dsigma);
  T* dR = static_cast<T*>(L.dR);
  T* df = static_cast<T*>(L.df);
  // lane r < 15 stores accumulator r
  T mine = acc[0];
#pragma unroll
  for (int i = 1; i < 15; ++i) mine = (r == i) ? acc[i] : mine;
  if (r < 3) {
    if (dp) dp[row * 3 + r] = mine;
  } else if (r < 6) {
    if (ds) ds[row * 3 + (r - 3)] = mine;
  } else if (r < 15) {
    if (dR) dR[row * 9 + (r - 6)] = mine;
  }
  if (df) {
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
      if (ft < F) {
        PVec<T, VEC> x;
#pragma unroll
        for (int i = 0; i < VEC; ++i) x.v[i] = dfa[c][i];
        *reinterpret_cast<PVec<T, VEC>*>(df + row * F + ft) = x;
      }
    }
  }
}

// ------------------------------------------------------------------ dispatch
template <typename T>
static int pick_vec_points(const PointLaunch& L) {
  const int maxv = 16 / static_cast<int>(sizeof(T));
  auto aligned = [&](const void* ptr, int v) {
    return ptr == nullptr || (reinterpret_cast<uintptr_t>(ptr) % (sizeof(T) * v)) == 0;
  };
  // no wider than one chunk needs: 16 lanes x v >= F with v as small as that allows
  int v0 = 1;
  while (v0 < maxv && 16 * v0 < L.F) v0 <<= 1;
  for (int v = v0; v > 1; v >>= 1)
    if (L.F % v == 0 && aligned(L.f, v) && aligned(L.df, v) && aligned(L.grad_f_out, v)) return v;
  return 1;
}

constexpr int kPointGroups = 4;  // 16-lane groups per point row

// The point kernels always weight with the anisotropic D (the key mode only ranks), so
// they are instantiated once for both key modes.
template <typename T, int VEC, bool SECOND, bool FULL, bool BC>
static hipError_t launch_p_ch(const PointLaunch& L, hipStream_t st) {
  constexpr int G = kPointGroups;
  constexpr int KEY = AKNN_KEY_ANISO;
  const int64_t per = 16 * VEC;
  const int64_t ch = (L.F + per - 1) / per;
  const int64_t rows_per_block = 16 / G;
  const dim3 grid(static_cast<unsigned>((L.rows + rows_per_block - 1) / rows_per_block)), block(256);
  if constexpr (FULL) {
    if (ch <= 1) hipLaunchKernelGGL((point_kernel_full<T, VEC, 1, KEY, true, G, true>), grid, block, 0, st, L);
    else if (ch <= 2) hipLaunchKernelGGL((point_kernel_full<T, VEC, 2, KEY, true, G, true>), grid, block, 0, st, L);
    else if (ch <= 4) hipLaunchKernelGGL((point_kernel_full<T, VEC, 4, KEY, true, G, true>), grid, block, 0, st, L);
    else if (ch <= 8) hipLaunchKernelGGL((point_kernel_full<T, VEC, 8, KEY, true, G, true>), grid, block, 0, st, L);
    else return hipErrorNotSupported;
  } else {
    if (ch <= 1) hipLaunchKernelGGL((point_kernel<T, VEC, 1, KEY, SECOND, G, BC>), grid, block, 0, st, L);
    else if (ch <= 2) hipLaunchKernelGGL((point_kernel<T, VEC, 2, KEY, SECOND, G, BC>), grid, block, 0, st, L);
    else if (ch <= 4) hipLaunchKernelGGL((point_kernel<T, VEC, 4, KEY, SECOND, G, BC>), grid, block, 0, st, L);
    else if (ch <= 8) hipLaunchKernelGGL((point_kernel<T, VEC, 8, KEY, SECOND, G, BC>), grid, block, 0, st, L);
    else return hipErrorNotSupported;
  }
  return hipGetLastError();
}

template <typename T, bool SECOND, bool FULL, bool BC>
static hipError_t launch_p_vf(const PointLaunch& L, hipStream_t st) {
  const int v = pick_vec_points<T>(L);
  if constexpr (sizeof(T) == 4) {
    if (v == 4) return launch_p_ch<T, 4, SECOND, FULL, BC>(L, st);
  }
  if (v == 2) return launch_p_ch<T, 2, SECOND, FULL, BC>(L, st);
  return launch_p_ch<T, 1, SECOND, FULL, BC>(L, st);
}

template <typename T, bool SECOND>
static hipError_t launch_p_vec(const PointLaunch& L, hipStream_t st) {
  if constexpr (SECOND) {
    if (L.gg_p || L.gg_f || L.gg_sigma || L.gg_R) {
      if (L.g_bcast) return hipErrorInvalidValue;  // the full second order takes a materialised upstream
      return launch_p_vf<T, true, true, false>(L, st);
    }
  }
  if (L.g_bcast) return launch_p_vf<T, SECOND, false, true>(L, st);
  return launch_p_vf<T, SECOND, false, false>(L, st);
}

template <bool SECOND>
static hipError_t launch_p(const PointLaunch& L, hipStream_t st) {
  if (L.dtype == AKNN_F32) return launch_p_vec<float, SECOND>(L, st);
  return launch_p_vec<double, SECOND>(L, st);
}

hipError_t launch_point_grad(const PointLaunch& L, hipStream_t st) {
  if (L.rows == 0) return hipSuccess;
  return L.second ? launch_p<true>(L, st) : launch_p<false>(L, st);
}

// ------------------------------------------------------------------ point dot products
// Dot records of a broadcast upstream: for every point row n the 16 values
// {p0 p1 p2, R00..R22, s0 s1 s2, a} with a = grad_f_out[cloud(n)] . f[n].  16 lanes per
// row in the point kernels' feature layout and reduction, so a has their bits; lane r
// then writes value r of the record (one coalesced 16-value row).  dq_dots reads a
// neighbour's geometry and dot in one record.
template <typename T, int VEC, int CH>
__global__ __launch_bounds__(256) void point_dots_kernel(PointLaunch L, T* __restrict__ out) {
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const T* __restrict__ g = static_cast<const T*>(L.grad_f_out);
  const int64_t F = L.F;
  const int r = threadIdx.x & 15;
  const int64_t gi = static_cast<int64_t>(blockIdx.x) * 16 + (threadIdx.x >> 4);
  const bool row_ok = gi < L.rows;
  const int64_t row = row_ok ? gi : L.rows - 1;
  T fsl[CH][VEC], gbc[CH][VEC];
  int64_t fofs[CH];
  bool fok[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int64_t ft = static_cast<int64_t>(c * 16 + r) * VEC;
    fok[c] = ft < F;
    fofs[c] = fok[c] ? ft : 0;
#pragma unroll
    for (int i = 0; i < VEC; ++i) fsl[c][i] = T(0);
    if (ft < F) {
      const PVec<T, VEC> x = *reinterpret_cast<const PVec<T, VEC>*>(f + row * F + ft);
#pragma unroll
      for (int i = 0; i < VEC; ++i) fsl[c][i] = x.v[i];
    }
  }
  const T a = bcast_row_dot<T, VEC, CH>(g, L.cl.cloud_of_point(row), F, fofs, fok, fsl, gbc, r);
  T val;
  if (r < 3) val = static_cast<const T*>(L.p)[row * 3 + r];
  else if (r < 12) val = static_cast<const T*>(L.R)[row * 9 + (r - 3)];
  else if (r < 15) val = static_cast<const T*>(L.sigma)[row * 3 + (r - 12)];
  else val = a;
  if (row_ok) out[row * 16 + r] = val;
}

template <typename T, int VEC>
static hipError_t launch_dots_ch(const PointLaunch& L, void* a, hipStream_t st) {
  const int64_t per = 16 * VEC;
  const int64_t ch = (L.F + per - 1) / per;
  const dim3 grid(static_cast<unsigned>((L.rows + 15) / 16)), block(256);
  T* out = static_cast<T*>(a);
  if (ch <= 1) hipLaunchKernelGGL((point_dots_kernel<T, VEC, 1>), grid, block, 0, st, L, out);
  else if (ch <= 2) hipLaunchKernelGGL((point_dots_kernel<T, VEC, 2>), grid, block, 0, st, L, out);
  else if (ch <= 4) hipLaunchKernelGGL((point_dots_kernel<T, VEC, 4>), grid, block, 0, st, L, out);
  else if (ch <= 8) hipLaunchKernelGGL((point_dots_kernel<T, VEC, 8>), grid, block, 0, st, L, out);
  else return hipErrorNotSupported;
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_dots_t(const PointLaunch& L, void* a, hipStream_t st) {
  const int v = pick_vec_points<T>(L);  // the point kernels' choice for the same pointers
  if constexpr (sizeof(T) == 4) {
    if (v == 4) return launch_dots_ch<T, 4>(L, a, st);
  }
  if (v == 2) return launch_dots_ch<T, 2>(L, a, st);
  return launch_dots_ch<T, 1>(L, a, st);
}

hipError_t launch_point_dots(const PointLaunch& L, void* a, hipStream_t st) {
  if (L.rows == 0) return hipSuccess;
  return L.dtype == AKNN_F32 ? launch_dots_t<float>(L, a, st) : launch_dots_t<double>(L, a, st);
}

// ------------------------------------------------------------------ dq from contributions
// dq[q] = sum over k of contrib[q*K + k] in slot order (fixed association).
// dq[q] = sum over k of entry (q, k)'s share, in slot order: one thread per query.  When
// a query's 3K shares come in whole 16-byte vectors (FP32, K a multiple of 4) the
// thread reads them as such, otherwise one value at a time.
// K = 16, FP32: each wave stages its 64 queries' 12 KB of shares through LDS with
// coalesced 16-byte loads, then each lane sums its own query in the same order as
// contrib_reduce_kernel (same bits).
__global__ __launch_bounds__(256) void contrib_reduce16_kernel(const float4* __restrict__ contrib, int64_t rows,
                                                               float* __restrict__ dq) {
  constexpr int kPer = 12;  // float4 per query: 16 entries x 3 values
  __shared__ float4 buf[4][64 * kPer];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * 256 + wave * 64;
  const int64_t n4 = rows * kPer;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int64_t g = q0 * kPer + j * 64 + lane;
    if (g < n4) buf[wave][j * 64 + lane] = contrib[g];
  }
  __syncthreads();
  const int64_t q = q0 + lane;
  if (q >= rows) return;
  const float4* v = buf[wave] + lane * kPer;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
  for (int k = 0; k < kPer; k += 3) {  // four entries = three vectors
    const float4 a = v[k], b = v[k + 1], c = v[k + 2];
    s0 += a.x; s1 += a.y; s2 += a.z;
    s0 += a.w; s1 += b.x; s2 += b.y;
    s0 += b.z; s1 += b.w; s2 += c.x;
    s0 += c.y; s1 += c.z; s2 += c.w;
  }
  dq[q * 3 + 0] = s0;
  dq[q * 3 + 1] = s1;
  dq[q * 3 + 2] = s2;
}

template <typename T, bool VEC4>
__global__ __launch_bounds__(256) void contrib_reduce_kernel(const T* __restrict__ contrib, int64_t rows, int K,
                                                             T* __restrict__ dq) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= rows) return;
  const T* src = contrib + q * K * 3;
  T s0 = T(0), s1 = T(0), s2 = T(0);
  if constexpr (VEC4) {
    const float4* v = reinterpret_cast<const float4*>(src);
    for (int k = 0; k < K; k += 4) {  // four entries = three vectors
      const float4 a = v[0], b = v[1], c = v[2];
      v += 3;
      s0 += a.x; s1 += a.y; s2 += a.z;
      s0 += a.w; s1 += b.x; s2 += b.y;
      s0 += b.z; s1 += b.w; s2 += c.x;
      s0 += c.y; s1 += c.z; s2 += c.w;
    }
  } else {
    for (int k = 0; k < K; ++k) {
      s0 += src[k * 3 + 0];
      s1 += src[k * 3 + 1];
      s2 += src[k * 3 + 2];
    }
  }
  dq[q * 3 + 0] = s0;
  dq[q * 3 + 1] = s1;
  dq[q * 3 + 2] = s2;
}

hipError_t launch_contrib_reduce(const void* contrib, int64_t rows, int64_t K, void* dq, int32_t dtype,
                                 hipStream_t st) {
  if (rows == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((rows + 255) / 256)), block(256);
  const bool vec4 = dtype == AKNN_F32 && K % 4 == 0 && (reinterpret_cast<uintptr_t>(contrib) % 16) == 0;
  if (dtype == AKNN_F32 && vec4 && K == 16)
    hipLaunchKernelGGL(contrib_reduce16_kernel, grid, block, 0, st, static_cast<const float4*>(contrib), rows,
                       static_cast<float*>(dq));
  else if (dtype == AKNN_F32 && vec4)
    hipLaunchKernelGGL((contrib_reduce_kernel<float, true>), grid, block, 0, st,
                       static_cast<const float*>(contrib), rows, static_cast<int>(K), static_cast<float*>(dq));
  else if (dtype == AKNN_F32)
    hipLaunchKernelGGL((contrib_reduce_kernel<float, false>), grid, block, 0, st,
                       static_cast<const float*>(contrib), rows, static_cast<int>(K), static_cast<float*>(dq));
  else
    hipLaunchKernelGGL((contrib_reduce_kernel<double, false>), grid, block, 0, st,
                       static_cast<const double*>(contrib), rows, static_cast<int>(K), static_cast<double*>(dq));
  return hipGetLastError();
}

// ------------------------------------------------------------------ CSR transpose
// keys[e] = global point row of entry e = q*K + k.
__global__ void csr_keys_kernel(const int64_t* __restrict__ idxs, Clouds cl, int64_t K,
                                int64_t entries, uint32_t* __restrict__ keys) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  const int64_t qrow = e / K;
  const int64_t b = cl.cloud_of_query(qrow);
  keys[e] = static_cast<uint32_t>(cl.p_begin(b) + idxs[e]);
}

// rank[order[i]] = i: a query row's position in the spatial order.
__global__ void rank_kernel(const int32_t* __restrict__ order, int64_t n, int32_t* __restrict__ rank) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) rank[order[i]] = static_cast<int32_t>(i);
}

// Four entries per thread (K divisible by 4, so the four share a query row): one
// 32-bit division and one cloud lookup for the four, 16-byte loads and store.  The
// entry count is below 2^31 (the argument check), and idxs is 16-byte aligned.
__global__ void csr_keys4_kernel(const int64_t* __restrict__ idxs, Clouds cl, uint32_t K,
                                 uint32_t entries, uint32_t* __restrict__ keys) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t e = 4 * t;
  if (e >= entries) return;
  const int64_t qrow = FastDiv(K).div(e);
  const uint32_t pb = static_cast<uint32_t>(cl.p_begin(cl.cloud_of_query(qrow)));
  const longlong2 a = *reinterpret_cast<const longlong2*>(idxs + e);
  const longlong2 c = *reinterpret_cast<const longlong2*>(idxs + e + 2);
  *reinterpret_cast<uint4*>(keys + e) = make_uint4(pb + static_cast<uint32_t>(a.x), pb + static_cast<uint32_t>(a.y),
                                                   pb + static_cast<uint32_t>(c.x), pb + static_cast<uint32_t>(c.y));
}

// off[i] = first position of point row i in the sorted keys (lower bound), i in [0, rows].
__global__ void csr_offsets_kernel(const uint32_t* __restrict__ sorted, int64_t entries,
                                   int64_t rows, int32_t* __restrict__ off) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > rows) return;
  int64_t lo = 0, hi = entries;
  const uint32_t key = static_cast<uint32_t>(i);
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < key) lo = mid + 1; else hi = mid;
  }
  off[i] = static_cast<int32_t>(lo);
}

static int key_bits(int64_t rows) {
  int bits = 1;
  while ((int64_t(1) << bits) < rows) ++bits;
  return bits;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Keys in and out and the pair sort's temporaries, or the packed sort's workspace.
static size_t csr_sort_bytes(int64_t entries) {
  const size_t pairs = 2 * align256(static_cast<size_t>(entries) * 4) + align256(radix_sort_workspace(entries, 4));
  const size_t packed = align256(csr_packed_workspace(entries));
  return pairs > packed ? pairs : packed;
}

hipError_t csr_temp_bytes(int64_t entries, int64_t point_rows, int64_t query_rows, size_t* bytes) {
  (void)point_rows;
  // the sort, then the query ranks of an ordered transpose
  *bytes = csr_sort_bytes(entries) + align256(static_cast<size_t>(query_rows) * 4);
  return hipSuccess;
}

hipError_t launch_build_csr(const int64_t* idxs, Clouds cl, int64_t query_rows, int64_t K,
                            int64_t point_rows, int32_t* csr_off, int32_t* csr_ent, void* ws,
                            size_t ws_bytes, hipStream_t st, const int32_t* qorder) {
  (void)ws_bytes;
  const int64_t entries = query_rows * K;
  char* base = static_cast<char*>(ws);
  uint32_t* keys_in = reinterpret_cast<uint32_t*>(base);
  uint32_t* keys_out = reinterpret_cast<uint32_t*>(base + align256(entries * 4));
  void* tmp = base + 2 * align256(entries * 4);
  // in the spatial query order: the sort's values (entry numbers, which keep each row's
  // entries in ascending q*K+k order) become position*K + k through the rank array
  int32_t* rank = nullptr;
  if (qorder && entries > 0) {
    rank = reinterpret_cast<int32_t*>(base + csr_sort_bytes(entries));
    hipLaunchKernelGGL(rank_kernel, dim3(static_cast<unsigned>((query_rows + 255) / 256)), dim3(256), 0, st,
                       qorder, query_rows, rank);
  }
  if (csr_packed_fits(entries, K, key_bits(point_rows)) && reinterpret_cast<uintptr_t>(idxs) % 16 == 0)
    return csr_sort_packed(idxs, cl, K, entries, point_rows, rank, csr_off, reinterpret_cast<uint32_t*>(csr_ent), ws,
                           st);
  if (entries > 0) {
    const int bs = 256;
    if (K % 4 == 0 && reinterpret_cast<uintptr_t>(idxs) % 16 == 0) {
      const int64_t threads = entries / 4;
      hipLaunchKernelGGL(csr_keys4_kernel, dim3(static_cast<unsigned>((threads + bs - 1) / bs)), dim3(bs), 0, st,
                         idxs, cl, static_cast<uint32_t>(K), static_cast<uint32_t>(entries), keys_in);
    } else {
      hipLaunchKernelGGL(csr_keys_kernel, dim3(static_cast<unsigned>((entries + bs - 1) / bs)), dim3(bs),
                         0, st, idxs, cl, K, entries, keys_in);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    err = radix_sort_pairs<uint32_t>(keys_in, keys_out, nullptr, reinterpret_cast<uint32_t*>(csr_ent),
                                     entries, 0, key_bits(point_rows), tmp, st, rank, static_cast<uint32_t>(K));
    if (err != hipSuccess) return err;
  }
  const int bs = 256;
  hipLaunchKernelGGL(csr_offsets_kernel, dim3(static_cast<unsigned>((point_rows + 1 + bs - 1) / bs)),
                     dim3(bs), 0, st, keys_out, entries, point_rows, csr_off);
  return hipGetLastError();
}

template <typename T>
__global__ void gather3_kernel(const T* __restrict__ src, const int32_t* __restrict__ order, T* __restrict__ out,
                               int64_t rows) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int64_t r = order[i];
  out[i * 3 + 0] = src[r * 3 + 0];
  out[i * 3 + 1] = src[r * 3 + 1];
  out[i * 3 + 2] = src[r * 3 + 2];
}

hipError_t launch_gather3(const void* src, const int32_t* order, void* out, int64_t rows, int32_t dtype,
                          hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((rows + 255) / 256)), block(256);
  if (dtype == AKNN_F32)
    hipLaunchKernelGGL((gather3_kernel<float>), grid, block, 0, st, static_cast<const float*>(src), order,
                       static_cast<float*>(out), rows);
  else
    hipLaunchKernelGGL((gather3_kernel<double>), grid, block, 0, st, static_cast<const double*>(src), order,
                       static_cast<double*>(out), rows);
  return hipGetLastError();
}

}  // namespace aknn
