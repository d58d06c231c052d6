// Modified Bessel functions K0 and K1 in FP64 for the modified-Helmholtz
// kernel K0(k r)/(2 pi) and its gradient -k K1(k r)/(2 pi r) (t - s)
// (examples/pure_kifmm_modified_helmholtz_gradient.py:25-68).
//
//   x <= 2: the ascending series (Abramowitz & Stegun 9.6.13, 9.6.11)
//     K0(x) = -(ln(x/2) + gamma) I0(x) + sum_{k>=1} H_k q^k / (k!)^2,
//     K1(x) = 1/x + ln(x/2) I1(x)
//             - (x/4) sum_{k>=0} (psi(k+1) + psi(k+2)) q^k / (k! (k+1)!),
//     q = x^2/4, I0 = sum q^k/(k!)^2, I1 = (x/2) sum q^k/(k!(k+1)!),
//     psi(k+1) = H_k - gamma; 14 terms reach the FP64 rounding level.  The
//     four sums are polynomials in q with coefficients fixed at compile time,
//     evaluated by Horner's rule (one FMA per term each).
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

// Coefficients of the ascending series in q = x^2/4: i0 = 1/(k!)^2,
// s0 = H_k/(k!)^2, i1 = 1/(k!(k+1)!), s1 = (H_k + H_{k+1} - 2 gamma)/(k!(k+1)!).
struct BesselSeries {
  double i0[kBesselSeriesTerms], s0[kBesselSeriesTerms];
  double i1[kBesselSeriesTerms], s1[kBesselSeriesTerms];
};
constexpr BesselSeries make_bessel_series() {
  BesselSeries c{};
  double f = 1.0, g = 1.0, h = 0.0;   // 1/(k!)^2, 1/(k!(k+1)!), H_k
  for (int k = 0; k < kBesselSeriesTerms; ++k) {
    if (k > 0) {
      f /= (double)k * (double)k;
      g /= (double)k * (double)(k + 1);
      h += 1.0 / (double)k;
    }
    const double h1 = h + 1.0 / (double)(k + 1);
    c.i0[k] = f;
    c.s0[k] = h * f;
    c.i1[k] = g;
    c.s1[k] = (h + h1 - 2.0 * kEulerGamma) * g;
  }
  return c;
}
constexpr BesselSeries kBesselSeries = make_bessel_series();
// K0 alone reads the coefficients from constant memory on the device (scalar
// loads), so that the FP64 constants do not occupy vector registers across
// the pair loops of the potential kernels (measured faster there).
__constant__ BesselSeries c_bessel_series = make_bessel_series();
__host__ __device__ inline const BesselSeries& bessel_series() {
#if defined(__HIP_DEVICE_COMPILE__)
  return c_bessel_series;
#else
  return kBesselSeries;
#endif
}

// x > 2 only
__host__ __device__ inline double bessel_k0_large(double x) {
  const double u = 4.0 / x - 1.0;
  return exp(-x) / sqrt(x) * cheb_sum(kK0Cheb, u);
}
__host__ __device__ inline void bessel_k01_large(double x, double& k0, double& k1) {
  const double u = 4.0 / x - 1.0;
  const double e = exp(-x) / sqrt(x);
  k0 = e * cheb_sum(kK0Cheb, u);
  k1 = e * cheb_sum(kK1Cheb, u);
}

__host__ __device__ inline double bessel_k0(double x) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    const BesselSeries& c = bessel_series();
    constexpr int n = kBesselSeriesTerms - 1;
    double i0 = c.i0[n], s = c.s0[n];
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      i0 = __builtin_fma(i0, q, c.i0[k]);
      s = __builtin_fma(s, q, c.s0[k]);
    }
    return s - (log(0.5 * x) + kEulerGamma) * i0;
  }
  return bessel_k0_large(x);
}

// K0 and K1 together (the gradient evaluator needs both).
__host__ __device__ inline void bessel_k01(double x, double& k0, double& k1) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    const double lg = log(0.5 * x);
    // compile-time coefficients here (the host path; the device kernels use
    // their own K0/K1 from d^2)
    const BesselSeries& c = kBesselSeries;
    constexpr int n = kBesselSeriesTerms - 1;
    double i0 = c.i0[n], s0 = c.s0[n], i1 = c.i1[n], s1 = c.s1[n];
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      i0 = __builtin_fma(i0, q, c.i0[k]);
      s0 = __builtin_fma(s0, q, c.s0[k]);
      i1 = __builtin_fma(i1, q, c.i1[k]);
      s1 = __builtin_fma(s1, q, c.s1[k]);
    }
    k0 = s0 - (lg + kEulerGamma) * i0;
    k1 = 1.0 / x + lg * (0.5 * x * i1) - 0.25 * x * s1;
    return;
  }
  bessel_k01_large(x, k0, k1);
}

}  // namespace fmm
