// FP64 scalar functions of the evidence path, compiled identically for the
// host driver and the gfx950 kernels (both with -ffp-contract=off), so that
// every evidence value is the same double wherever it is computed.
//
// Reference semantics (isakruas/fbls src/search/math.rs):
//   erfc                        math.rs:8-100 (W. J. Cody's rational
//                               approximations; Math. Comp. 23, 1969 --
//                               the published coefficient table)
//   log_add                     math.rs:103-110
//   positive_amplitude_evidence math.rs:115-140
//   subspace_upper              math.rs:145-177
//
// exp and log are this library's own (the reference uses the platform libm
// on the host and libdevice in kernels):
//   exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, ln2 split into a
//        40-bit head (k * head exact for |k| < 4096) and a tail; e^r - 1 by
//        its Taylor series to r^13 (remainder < 4e-18 relative); 2^k by
//        ldexp.  Faithful (within 1 ulp of the correctly rounded value).
//   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); f = m - 1 is exact;
//        s = f / (2 + f); log(1 + f) = f - (f^2/2 - s (f^2/2 + R)) with
//        R = sum_{i=1..11} 2 s^(2i) / (2i + 1), from 2 atanh(s) = 2s + s R and
//        2s = f - s f.  Faithful.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace fbls {
namespace fm {

constexpr double kInfinity = __builtin_inf();
constexpr double kLn2Hi = 0x1.62e42fefa2000p-1;
constexpr double kLn2Lo = 0x1.9ef35793c7673p-41;

__host__ __device__ inline double exp(double x) {
  if (x != x) return x + x;
  if (x > 709.782712893383973096) return kInfinity;
  if (x < -745.133219101941108420) return 0.0;
  const double kf = rint(x * 1.4426950408889634);
  const double r = (x - kf * kLn2Hi) - kf * kLn2Lo;
  double q = 1.0 / 6227020800.0;  // 1/13!
  q = q * r + 1.0 / 479001600.0;
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;
  const double em1 = r + (r * r) * q;
  return ldexp(1.0 + em1, (int)kf);
}

__host__ __device__ inline double log(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -kInfinity;
  if (x == kInfinity) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654757) {
    m = m * 2.0;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 23.0;
  R = R * z + 2.0 / 21.0;
  R = R * z + 2.0 / 19.0;
  R = R * z + 2.0 / 17.0;
  R = R * z + 2.0 / 15.0;
  R = R * z + 2.0 / 13.0;
  R = R * z + 2.0 / 11.0;
  R = R * z + 2.0 / 9.0;
  R = R * z + 2.0 / 7.0;
  R = R * z + 2.0 / 5.0;
  R = R * z + 2.0 / 3.0;
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)e;
  return dk * kLn2Hi - ((hfsq - (s * (hfsq + R) + dk * kLn2Lo)) - f);
}

__host__ __device__ inline double erfc(double x) {
  // Cody's coefficients (published table)
  const double A0 = 3.16112374387056560e0, A1 = 1.13864154151050156e2,
               A2 = 3.77485237685302021e2, A3 = 3.20937758913846947e3,
               A4 = 1.85777706184603153e-1;
  const double B0 = 2.36012909523441209e1, B1 = 2.44024637934444173e2,
               B2 = 1.28261652607737228e3, B3 = 2.84423683343917062e3;
  const double C[9] = {5.64188496988670089e-1, 8.88314979438837594e0,
                       6.61191906371416295e1,  2.98635138197400131e2,
                       8.81952221241769090e2,  1.71204761263407058e3,
                       2.05107837782607147e3,  1.23033935479799725e3,
                       2.15311535474403846e-8};
  const double D[8] = {1.57449261107098347e1, 1.17693950891312499e2,
                       5.37181101862009858e2, 1.62138957456669019e3,
                       3.29079923573345963e3, 4.36261909014324716e3,
                       3.43936767414372164e3, 1.23033935480374942e3};
  const double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1,
                       1.25781726111229246e-1, 1.60837851487422766e-2,
                       6.58749161529837803e-4, 1.63153871373020978e-2};
  const double Q[5] = {2.56852019228982242e0, 1.87295284992346725e0,
                       5.27905102951428412e-1, 6.05183413124413191e-2,
                       2.33520497626869185e-3};
  const double y = fabs(x);
  double result;
  if (y <= 0.46875) {
    const double ysq = y > 1.11e-16 ? y * y : 0.0;
    double xnum = A4 * ysq, xden = ysq;
    xnum = (xnum + A0) * ysq;
    xden = (xden + B0) * ysq;
    xnum = (xnum + A1) * ysq;
    xden = (xden + B1) * ysq;
    xnum = (xnum + A2) * ysq;
    xden = (xden + B2) * ysq;
    return 1.0 - x * (xnum + A3) / (xden + B3);
  } else if (y <= 4.0) {
    double xnum = C[8] * y, xden = y;
    for (int i = 0; i < 7; i++) {
      xnum = (xnum + C[i]) * y;
      xden = (xden + D[i]) * y;
    }
    const double r = (xnum + C[7]) / (xden + D[7]);
    const double ysq = trunc(y * 16.0) / 16.0;
    const double del = (y - ysq) * (y + ysq);
    result = fm::exp(-ysq * ysq) * fm::exp(-del) * r;
  } else if (y < 26.543) {
    const double ysq = 1.0 / (y * y);
    double xnum = P[5] * ysq, xden = ysq;
    for (int i = 0; i < 4; i++) {
      xnum = (xnum + P[i]) * ysq;
      xden = (xden + Q[i]) * ysq;
    }
    double r = ysq * (xnum + P[4]) / (xden + Q[4]);
    r = (5.6418958354775628695e-1 - r) / y;
    const double y16 = trunc(y * 16.0) / 16.0;
    const double del = (y - y16) * (y + y16);
    result = fm::exp(-y16 * y16) * fm::exp(-del) * r;
  } else {
    result = 0.0;
  }
  return x < 0.0 ? 2.0 - result : result;
}

__host__ __device__ inline double log_add(double x, double y) {
  const double high = fmax(x, y);
  if (high == -kInfinity) return high;
  return high + fm::log(1.0 + fm::exp(fmin(x, y) - high));
}

// log of the integral over depth >= 0 of a half-normal prior of scale tau
// times exp(a d - q d^2 / 2), relative to d = 0
__host__ __device__ inline double positive_amplitude_evidence(double a, double q,
                                                              double tau) {
  q = fmax(q, 0.0);
  const double precision = 1.0 + tau * tau * q;
  const double z = tau * a / sqrt(precision);
  if (z < -26.0) {
    const double zz = fmax(z * z, 1.0);
    const double inv = 1.0 / zz;
    const double series =
        1.0 + inv * (-1.0 + inv * (3.0 + inv * (-15.0 + inv * (105.0 +
              inv * (-945.0 + inv * (10395.0 - inv * 135135.0))))));
    return -0.5 * fm::log(precision) - fm::log(fmax(-z, 1.0)) -
           0.22579135264472743236 + fm::log(series);
  }
  return -0.5 * fm::log(precision) + 0.5 * z * z +
         fm::log(fmax(fm::erfc(-z * 0.70710678118654752440), 1e-300));
}

// upper bound of the log evidence of every positive mixture of two
// templates (numerators a, b; precisions q, r; cross term)
__host__ __device__ inline double subspace_upper(double a, double b, double q,
                                                 double r, double cross,
                                                 double rho2) {
  const double product = q * r;
  const double det = q * r - cross * cross;
  const double energy =
      (a * a * r + b * b * q - 2.0 * a * b * cross) / fmax(det * rho2, 1e-300);
  const bool safe = q > 0.0 && r > 0.0 && det > 1e-8 * product;
  const double pa = fmax(a, 0.0), pb = fmax(b, 0.0);
  const double boundary = 0.5 * fmax(pa * pa / fmax(q * rho2, 1e-300),
                                     pb * pb / fmax(r * rho2, 1e-300));
  const bool interior = a * r - b * cross > 0.0 && b * q - a * cross > 0.0;
  double upper = safe ? (interior ? 0.5 * fmax(energy, 0.0) : boundary) : kInfinity;
  const bool only_a = r == 0.0 && b == 0.0 && cross == 0.0;
  const bool only_b = q == 0.0 && a == 0.0 && cross == 0.0;
  if (only_a && q > 0.0) upper = 0.5 * a * a / (q * rho2);
  if (only_b && r > 0.0) upper = 0.5 * b * b / (r * rho2);
  if (only_a && only_b) upper = 0.0;
  if (a <= 0.0 && b <= 0.0) upper = 0.0;
  return upper * 1.000001 + 1e-6;
}

}  // namespace fm
}  // namespace fbls
