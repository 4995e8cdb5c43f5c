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
template <typename T>
AKNN_HD inline T box_lower_bound(float lx, float ly, float lz, float hx, float hy, float hz, float mu, T qx, T qy,
                                 T qz) {
  const T dx = fmax(fmax(static_cast<T>(lx) - qx, qx - static_cast<T>(hx)), T(0));
  const T dy = fmax(fmax(static_cast<T>(ly) - qy, qy - static_cast<T>(hy)), T(0));
  const T dz = fmax(fmax(static_cast<T>(lz) - qz, qz - static_cast<T>(hz)), T(0));
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
