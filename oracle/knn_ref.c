/*
 * Sequential-selection restatement of the reference neighbour search -- TEST
 * INFRASTRUCTURE ONLY (loaded by the test suite through ctypes; never by the
 * shipped library).
 *
 * What it restates:
 *   - the key, csrc/forward.cu:85-92 (anisotropic) and csrc/forward.cu:33-36 / :154
 *     (Euclidean), one rounding per operation in the reference's association order;
 *     build with -ffp-contract=off so no multiply-add is fused;
 *   - the bounded "keep the K smallest" scan of csrc/mink.cuh:60-81 followed by the
 *     ordering pass of csrc/mink.cuh:99-112,
 * with one deliberate difference: candidates are ordered by the total order on
 * (key, point index), so exact key ties keep the lower index.
 *
 * It is an algorithm independent of the stable-sort oracle in oracle/plain.py
 * (selection with max-eviction instead of sorting), so the two cross-check each other.
 */
#include <stdint.h>
#include <stdlib.h>

#define DEFINE_KNN(NAME, T)                                                              \
  static inline T NAME##_key(const T* qq, const T* pp, const T* ss, const T* RR,          \
                             int mode) {                                                  \
    T x = qq[0] - pp[0];                                                                  \
    T y = qq[1] - pp[1];                                                                  \
    T z = qq[2] - pp[2];                                                                  \
    if (mode == 1) {                                                                      \
      T e = x * x;                                                                        \
      e = e + y * y;                                                                      \
      e = e + z * z;                                                                      \
      return e;                                                                           \
    }                                                                                     \
    T j = RR[0] * x; j = j + RR[1] * y; j = j + RR[2] * z;                                \
    T k = RR[3] * x; k = k + RR[4] * y; k = k + RR[5] * z;                                \
    T l = RR[6] * x; l = l + RR[7] * y; l = l + RR[8] * z;                                \
    T D = (j * j) * ss[0];                                                                \
    D = D + (k * k) * ss[1];                                                              \
    D = D + (l * l) * ss[2];                                                              \
    return D;                                                                             \
  }                                                                                       \
  /* (a,ia) < (b,ib) in the (key, index) order */                                         \
  static inline int NAME##_less(T a, int64_t ia, T b, int64_t ib) {                       \
    return a < b || (a == b && ia < ib);                                                  \
  }                                                                                       \
  int NAME(const T* q, const T* p, const T* sigma, const T* R, int64_t B, int64_t Q,      \
           int64_t P, int64_t K, int mode, T* dists, int64_t* idxs) {                     \
    if (K < 1 || K > P) return -1;                                                        \
    for (int64_t b = 0; b < B; ++b) {                                                     \
      _Pragma("omp parallel for schedule(static)")                                        \
      for (int64_t iq = 0; iq < Q; ++iq) {                                                \
        T* kk = dists + (b * Q + iq) * K;                                                 \
        int64_t* vv = idxs + (b * Q + iq) * K;                                            \
        const T* qq = q + (b * Q + iq) * 3;                                               \
        int64_t size = 0, arg = 0;                                                        \
        for (int64_t n = 0; n < P; ++n) {                                                 \
          const int64_t pn = b * P + n;                                                   \
          T D = NAME##_key(qq, p + pn * 3, sigma + pn * 3, R + pn * 9, mode);             \
          if (size < K) {                                                                 \
            kk[size] = D; vv[size] = n; size++;                                           \
            if (size == K) { /* locate the largest (key, index) */                        \
              arg = 0;                                                                    \
              for (int64_t s = 1; s < K; ++s)                                             \
                if (NAME##_less(kk[arg], vv[arg], kk[s], vv[s])) arg = s;                 \
            }                                                                             \
          } else if (NAME##_less(D, n, kk[arg], vv[arg])) {                               \
            kk[arg] = D; vv[arg] = n;                                                     \
            arg = 0;                                                                      \
            for (int64_t s = 1; s < K; ++s)                                               \
              if (NAME##_less(kk[arg], vv[arg], kk[s], vv[s])) arg = s;                   \
          }                                                                               \
        }                                                                                 \
        /* insertion sort by (key, index) */                                              \
        for (int64_t s = 1; s < K; ++s) {                                                 \
          T ck = kk[s]; int64_t cv = vv[s]; int64_t t = s - 1;                            \
          while (t >= 0 && NAME##_less(ck, cv, kk[t], vv[t])) {                           \
            kk[t + 1] = kk[t]; vv[t + 1] = vv[t]; --t;                                    \
          }                                                                               \
          kk[t + 1] = ck; vv[t + 1] = cv;                                                 \
        }                                                                                 \
      }                                                                                   \
    }                                                                                     \
    return 0;                                                                             \
  }

DEFINE_KNN(knn_ref_f32, float)
DEFINE_KNN(knn_ref_f64, double)
