// Certified lower bounds of the neighbour key, for the exact tile pruning of the search.
//
// Compiled for the device by the search kernels and for the host by the CPU test of the
// certificate (the same source).  For a point n with M = R^T diag(s) R and
// lmin = lambda_min(M), the exact key D = d^T M d >= lmin |d|^2; the FP32 (or FP64)
// evaluation of the key with one rounding per operation (unit roundoff u, g = 4u/(1-4u))
// satisfies key >= (1 - g)(D - 2 g |d| sqrt(D trace M)), increasing in D where it
// matters, hence key >= mu_n |d|^2 with mu_n = (1 - g)(lmin - 2 g sqrt(lmin trace M)).
// A box of points stores min mu_n (scaled by 1 - 1e-6 and rounded down to float, which
// absorbs the rounding of the bound's own evaluation), and mu * dist^2(q, box) then
// bounds every key of the box's points from below.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define AKNN_HD __host__ __device__
#else
#define AKNN_HD
#endif

namespace aknn {

// Neighbouring floats (IEEE bit stepping).
AKNN_HD inline float bits_to_float(uint32_t b) { return __builtin_bit_cast(float, b); }
AKNN_HD inline uint32_t float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
AKNN_HD inline float next_down(float f) {
  if (f != f || f == -INFINITY) return f;
  if (f == 0.f) return bits_to_float(0x80000001u);
  const uint32_t b = float_to_bits(f);
  return bits_to_float(f > 0.f ? b - 1 : b + 1);
}
AKNN_HD inline float next_up(float f) { return -next_down(-f); }

// A lower bound of the smallest eigenvalue of the symmetric 3x3 matrix M (trace S >=
// ||M||_F), by cyclic Jacobi rotations in FP64: each rotation zeroes one off-diagonal
// pair (tan t of the rotation angle the root of smaller magnitude of t^2 + 2 cot(2 angle)
// t - 1 = 0), so the iterate converges to diag(eigenvalues).  By Weyl's inequality the
// eigenvalues of the final iterate lie within its off-diagonal Frobenius norm of its
// diagonal, and the rotations in floating point are backward stable (the final iterate
// is an exact similarity of M plus a perturbation of a few dozen ulp of ||M||, covered
// by the 1e-12 S margin).  Unlike the closed form through acos, whose error grows to
// sqrt(eps) of the spread when two eigenvalues nearly coincide, this stays within
// eps-level of ||M||.
AKNN_HD inline double lambda_min_lower(const double (&M)[3][3], double S) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = M[i][j];
  for (int sweep = 0; sweep < 10; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (!(off > 1e-40 * S * S)) break;
    for (int pair = 0; pair < 3; ++pair) {
      const int p = pair == 2 ? 1 : 0, q = pair == 0 ? 1 : 2, r = 3 - p - q;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double at = fabs(theta);
      // t = sign(theta) / (|theta| + sqrt(theta^2 + 1)), ~1/(2 theta) for huge theta
      const double t = (at > 1e150 ? 0.5 / at : 1.0 / (at + sqrt(at * at + 1.0))) * (theta < 0.0 ? -1.0 : 1.0);
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      const double arp = a[r][p], arq = a[r][q];
      a[p][p] -= t * apq;
      a[q][q] += t * apq;
      a[p][q] = a[q][p] = 0.0;
      a[r][p] = a[p][r] = c * arp - sn * arq;
      a[r][q] = a[q][r] = sn * arp + c * arq;
    }
  }
  const double off = sqrt(2.0 * (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2]));
  const double d = fmin(fmin(a[0][0], a[1][1]), a[2][2]);
  return d - off - 1e-12 * S;
}

// Certified lower bound mu_n of key(q, n) / |q - p_n|^2 for a point with rotation r9
// (row-major) and scales s3, for keys evaluated in T; negative means "no bound" (the
// box must always be scanned): negative or non-finite scales, non-finite rotations.
// EUCLID: the plain squared distance (its own rounding only).
template <typename T, bool EUCLID>
AKNN_HD inline double point_mu(const T* r9, const T* s3) {
  constexpr double u = sizeof(T) == 4 ? 5.9604644775390625e-08 : 1.1102230246251565e-16;
  constexpr double g = 4.0 * u / (1.0 - 4.0 * u);
  if (EUCLID) return (1.0 - g) * (1.0 - 2.0 * g * 1.7320508075688772);
  double R[9], s[3];
  for (int i = 0; i < 9; ++i) R[i] = static_cast<double>(r9[i]);
  for (int i = 0; i < 3; ++i) s[i] = static_cast<double>(s3[i]);
  for (int i = 0; i < 9; ++i)
    if (!__builtin_isfinite(R[i])) return -1.0;
  for (int i = 0; i < 3; ++i)
    if (!(s[i] >= 0.0) || !__builtin_isfinite(s[i])) return -1.0;
  // M = R^T diag(s) R (symmetric positive semidefinite 3x3)
  double M[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c)
      M[a][c] = s[0] * R[a] * R[c] + s[1] * R[3 + a] * R[3 + c] + s[2] * R[6 + a] * R[6 + c];
  const double S = M[0][0] + M[1][1] + M[2][2];  // trace >= ||M||_F for PSD M
  const double lmin = lambda_min_lower(M, S);
  if (!(lmin > 0.0)) return 0.0;
  const double mu = (1.0 - g) * (lmin - 2.0 * g * sqrt(lmin * S));
  return mu > 0.0 ? mu : 0.0;
}

// The bound as a box header stores it: min mu of the box's points scaled by 1 - 1e-6 and
// rounded down to float; -1 for a box that must always be scanned.
AKNN_HD inline float stored_mu(double mu, bool bad) {
  if (bad || !(mu >= 0.0)) return -1.f;
  const double m = mu * (1.0 - 1e-6);
  float muf = static_cast<float>(m);
  if (static_cast<double>(muf) > m) muf = next_down(muf);
  return muf;
}

// mu * dist^2(q, box), in the key's precision, for the box [lo, hi] (float corners that
// contain the exact coordinates); meaningful for mu >= 0.
// The offset along one axis is q - clamp(q, lo, hi): -(lo - q) below the box, q - hi
// above it (the same rounded magnitudes as lo - q and q - hi), 0 inside.  On the device
// the FP32 clamp is one median-of-three (the same value for lo <= hi).
template <typename T>
AKNN_HD inline T clamp_to_box(T q, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_fmed3f(q, lo, hi);
#endif
  return fmin(fmax(q, static_cast<T>(lo)), static_cast<T>(hi));
}
template <typename T>
AKNN_HD inline T box_lower_bound(float lx, float ly, float lz, float hx, float hy, float hz, float mu, T qx, T qy,
                                 T qz) {
  const T dx = qx - clamp_to_box<T>(qx, lx, hx);
  const T dy = qy - clamp_to_box<T>(qy, ly, hy);
  const T dz = qz - clamp_to_box<T>(qz, lz, hz);
  return static_cast<T>(mu) * (dx * dx + dy * dy + dz * dz);
}

// Whether a lane whose current K-th key is thr still needs the box (header h = {lo x y z,
// hi x y z, mu, -}): always for a negative mu, else unless the bound exceeds thr.
template <typename T>
AKNN_HD inline bool box_needed(const float* h, T qx, T qy, T qz, T thr, bool valid) {
  const T lb = box_lower_bound<T>(h[0], h[1], h[2], h[3], h[4], h[5], h[6], qx, qy, qz);
  return valid & ((static_cast<T>(h[6]) < T(0)) | !(lb > thr));
}

}  // namespace aknn
