// Modified Bessel functions K0 and K1 in FP64 for the modified-Helmholtz
// kernel K0(k r)/(2 pi) and its gradient -k K1(k r)/(2 pi r) (t - s)
// (examples/pure_kifmm_modified_helmholtz_gradient.py:25-68).
//
//   x <= 2: the ascending series (Abramowitz & Stegun 9.6.13, 9.6.11)
//     K0(x) = -(ln(x/2) + gamma) I0(x) + sum_{k>=1} H_k q^k / (k!)^2,
//     K1(x) = 1/x + ln(x/2) I1(x)
//             - (x/4) sum_{k>=0} (psi(k+1) + psi(k+2)) q^k / (k! (k+1)!),
//     q = x^2/4, I0 = sum q^k/(k!)^2, I1 = (x/2) sum q^k/(k!(k+1)!),
//     psi(k+1) = H_k - gamma; 14 terms reach the FP64 rounding level.
//   x > 2: K_n(x) = e^{-x} / sqrt(x) * g_n(u), u = 4/x - 1, with g_n the
//     degree-24 Chebyshev interpolant of sqrt(x) e^x K_n(x) (bessel_coef.h),
//     summed by Clenshaw's recurrence.
// Relative error against scipy.special.k0/k1 is below 1e-14 on (0, 1e3].
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "bessel_coef.h"

namespace fmm {

constexpr double kEulerGamma = 0.57721566490153286061;
constexpr int kBesselSeriesTerms = 14;

__host__ __device__ inline double cheb_sum(const double* c, double u) {
  double b1 = 0.0, b2 = 0.0;
  const double u2 = 2.0 * u;
#pragma unroll
  for (int k = kBesselChebDeg; k >= 1; --k) {
    const double b0 = c[k] + u2 * b1 - b2;
    b2 = b1;
    b1 = b0;
  }
  return c[0] + u * b1 - b2;
}

__host__ __device__ inline double bessel_k0(double x) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    double t = 1.0, i0 = 1.0, s = 0.0, h = 0.0;
#pragma unroll
    for (int k = 1; k < kBesselSeriesTerms; ++k) {
      t *= q * (1.0 / ((double)k * (double)k));
      h += 1.0 / (double)k;
      i0 += t;
      s += h * t;
    }
    return s - (log(0.5 * x) + kEulerGamma) * i0;
  }
  const double u = 4.0 / x - 1.0;
  return exp(-x) / sqrt(x) * cheb_sum(kK0Cheb, u);
}

// K0 and K1 together (the gradient evaluator needs both).
__host__ __device__ inline void bessel_k01(double x, double& k0, double& k1) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    const double lg = log(0.5 * x);
    double t = 1.0, i0 = 1.0, s0 = 0.0, h = 0.0;       // K0 series
    double v = 1.0, i1 = 1.0, s1 = 1.0 - 2.0 * kEulerGamma;   // K1 series, k = 0 term
    double h1 = 1.0;                                            // H_{k+1}
#pragma unroll
    for (int k = 1; k < kBesselSeriesTerms; ++k) {
      const double dk = (double)k;
      t *= q * (1.0 / (dk * dk));
      h += 1.0 / dk;
      i0 += t;
      s0 += h * t;
      v *= q * (1.0 / (dk * (dk + 1.0)));
      h1 += 1.0 / (dk + 1.0);
      i1 += v;
      s1 += (h + h1 - 2.0 * kEulerGamma) * v;
    }
    k0 = s0 - (lg + kEulerGamma) * i0;
    k1 = 1.0 / x + lg * (0.5 * x * i1) - 0.25 * x * s1;
    return;
  }
  const double u = 4.0 / x - 1.0;
  const double e = exp(-x) / sqrt(x);
  k0 = e * cheb_sum(kK0Cheb, u);
  k1 = e * cheb_sum(kK1Cheb, u);
}

}  // namespace fmm
