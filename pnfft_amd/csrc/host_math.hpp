// host_math.hpp -- FP64 plan-time math on the host: window shape parameters,
// Bessel I0, the cardinal B-spline at an arbitrary argument, the inverse
// window Fourier coefficients of matrix D, and the scalar window functions of
// the public API (psi, dpsi, ddpsi, phi_hat, inv_phi_hat).
#pragma once

#include <cmath>

#include "window_math.hpp"

namespace pnfft_amd {
namespace host {

constexpr double kPi = 3.14159265358979323846;

// Flag bits of api/pnfft.h:302-334 that pick the window.
constexpr unsigned kFlagGaussian = 1u << 13;
constexpr unsigned kFlagBspline = 1u << 14;
constexpr unsigned kFlagSincPow = 1u << 15;
constexpr unsigned kFlagBesselI0 = 1u << 16;
constexpr unsigned kFlagFkGaussT = 1u << 17;

// Window selection order of kernel/ndft-parallel.c:1650-1675.
inline int window_from_flags(unsigned flags) {
  if (flags & kFlagGaussian) return WIN_GAUSS;
  if (flags & kFlagBspline) return WIN_BSPLINE;
  if (flags & kFlagSincPow) return WIN_SINCPOW;
  if (flags & kFlagBesselI0) return WIN_BESSEL_I0;
  return WIN_KB;
}

// Shape parameter b of kernel/ndft-parallel.c:1009-1064 for sigma = n/N.
inline double window_b(int win, int m, double sigma) {
  switch (win) {
    case WIN_GAUSS: return (double(m) / kPi) * 2.0 * sigma / (2.0 * sigma - 1.0);
    case WIN_BSPLINE: return 0.0;
    case WIN_SINCPOW: return double(m) * (2.0 * sigma) / (2.0 * sigma - 1.0);
    default: return kPi * (2.0 - 1.0 / sigma);
  }
}

// I0(x) in FP64 (the reference calls GSL, kernel/bessel_i0.c:372-375).
// Written from the mathematics: for |x| <= 25 the Taylor series
// sum_k (x^2/4)^k / (k!)^2 evaluated from the small end (positive terms);
// beyond, I0(x) = e^x/sqrt(2 pi x) * sum_k c_k / x^k with
// c_k = ((2k-1)!!)^2 / (k! 8^k) (Abramowitz & Stegun 9.7.1), cut at the
// first term below 1e-17 or at the smallest term.
inline double bessel_i0(double x) {
  const double ax = std::fabs(x);
  if (ax <= 25.0) {
    const double q = 0.25 * ax * ax;
    // terms t_k = q^k/(k!)^2; sum the tail backwards for accuracy
    double terms[200];
    int K = 0;
    double t = 1.0;
    terms[K++] = t;
    for (int k = 1; k < 200; k++) {
      t *= q / (double(k) * double(k));
      terms[K++] = t;
      if (t < 1e-19 * terms[0] && t < 1e-19) break;
      if (k > 2 * ax + 20 && t < 1e-18) break;
    }
    double s = 0.0;
    for (int k = K - 1; k >= 0; k--) s += terms[k];
    return s;
  }
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 80; k++) {
    const double c = 2.0 * k - 1.0;
    const double next = term * c * c / (8.0 * k * ax);
    if (next > term) break;
    term = next;
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return std::exp(ax) / std::sqrt(2.0 * kPi * ax) * sum;
}

// I1(x) in FP64 (the reference calls GSL, kernel/bessel_i1.c:567-597): the
// series sum_k (x/2)^{2k+1} / (k! (k+1)!) for |x| <= 25, beyond it
// e^x/sqrt(2 pi x) sum_k (-1)^k prod_{j<=k} (4 - (2j-1)^2) / (k! (8x)^k)
// (Abramowitz & Stegun 9.7.1 with mu = 4); odd in x.
inline double bessel_i1(double x) {
  const double ax = std::fabs(x);
  double v;
  if (ax <= 25.0) {
    const double q = 0.25 * ax * ax;
    double t = 0.5 * ax, sum = t;
    for (int k = 1; k < 200; k++) {
      t *= q / (double(k) * double(k + 1));
      sum += t;
      if (t < 1e-18 * sum) break;
    }
    v = sum;
  } else {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 80; k++) {
      const double c = 2.0 * k - 1.0;
      const double next = -term * (4.0 - c * c) / (8.0 * k * ax);
      if (std::fabs(next) > std::fabs(term)) break;
      term = next;
      sum += term;
      if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
    }
    v = std::exp(ax) / std::sqrt(2.0 * kPi * ax) * sum;
  }
  return x < 0 ? -v : v;
}

// erf(x + i y) for x > 0 (Abramowitz & Stegun 7.1.29, relative error about
// 1e-16), the complex error function of the truncated Gaussian's Fourier
// coefficients (the reference calls libcerf's cerf, kernel/matrix_D.c:37-64).
inline void cerf(double x, double y, double* re, double* im) {
  const double e = std::exp(-x * x), c = std::cos(2.0 * x * y), sn = std::sin(2.0 * x * y);
  double sr = 0.0, si = 0.0;
  for (int k = 1; k <= 60; k++) {
    const double w = std::exp(-0.25 * k * k) / (k * k + 4.0 * x * x);
    const double ch = std::cosh(k * y), sh = std::sinh(k * y);
    sr += w * (2.0 * x - 2.0 * x * ch * c + k * sh * sn);
    si += w * (2.0 * x * ch * sn + k * sh * c);
  }
  *re = std::erf(x) + e / (2.0 * kPi * x) * (1.0 - c) + 2.0 / kPi * e * sr;
  *im = e / (2.0 * kPi * x) * sn + 2.0 / kPi * e * si;
}

// phi_hat of the Gaussian truncated to the stencil, e^{-b (pi k/n)^2} Re erf(m/sqrt b + i pi k sqrt b/n)
// (phi_hat_gauss_t, kernel/matrix_D.c:57-64).
inline double phi_hat_gauss_t(long k, long n, double b, int m) {
  const double sb = std::sqrt(b), a = kPi * double(k) / double(n);
  double re, im;
  cerf(double(m) / sb, kPi * double(k) * sb / double(n), &re, &im);
  return std::exp(-a * a * b) * re;
}

// Cardinal B-spline M_k(x) on (0, k), zero elsewhere (kernel/bspline.c:45-91).
inline double bspline(int k, double x) {
  if (!(x > 0.0 && x < double(k)) || k > 2 * kMaxM) return 0.0;
  double V[2 * kMaxM];
  double fl = std::floor(x);
  double u = x - fl;
  int d0 = int(fl);
  if (u == 0.0) {
    u = 1.0;
    d0 -= 1;
  }
  bspline_values<double>(k, u, V);
  return V[d0];
}

// 1/phi_hat_t(k) of kernel/matrix_D.c:30-140, 191-207 (factor 1/n cancels, :29).
inline double inv_phi_hat(int win, long k, long n, double b, int m) {
  switch (win) {
    case WIN_GAUSS: {
      const double a = kPi * double(k) / double(n);
      return std::exp(a * a * b);
    }
    case WIN_BSPLINE: return std::pow(sinc_ref<double>(double(k) * kPi / double(n)), -2.0 * m);
    case WIN_SINCPOW: {
      const double d = std::fabs(double(k) * b / double(n));
      return (d < m) ? 1.0 / bspline(2 * m, d + m) : 0.0;
    }
    case WIN_BESSEL_I0: {
      const double a = 2.0 * kPi * double(k) / double(n);
      const double d = b * b - a * a;
      const double r = std::sqrt(std::fabs(d));
      if (d < 0) return r / std::sin(m * r);
      return (d > 0) ? r / std::sinh(m * r) : 1.0 / m;
    }
    default: {
      const double a = 2.0 * kPi * double(k) / double(n);
      const double d = b * b - a * a;
      if (d < 0) return 0.0;
      return (d > 0) ? 1.0 / bessel_i0(m * std::sqrt(d)) : 1.0;
    }
  }
}

// phi_hat_t(k) (kernel/matrix_D.c:209-225).
inline double phi_hat(int win, long k, long n, double b, int m) {
  switch (win) {
    case WIN_GAUSS: {
      const double a = kPi * double(k) / double(n);
      return std::exp(-a * a * b);
    }
    case WIN_BSPLINE: return std::pow(sinc_ref<double>(double(k) * kPi / double(n)), 2.0 * m);
    case WIN_SINCPOW: return bspline(2 * m, std::fabs(double(k) * b / double(n)) + m);
    case WIN_BESSEL_I0: {
      const double a = 2.0 * kPi * double(k) / double(n);
      const double d = b * b - a * a;
      const double r = std::sqrt(std::fabs(d));
      if (d < 0) return std::sin(m * r) / r;
      return (d > 0) ? std::sinh(m * r) / r : double(m);
    }
    default: {
      const double a = 2.0 * kPi * double(k) / double(n);
      const double d = b * b - a * a;
      if (d < 0) return 0.0;
      return (d > 0) ? bessel_i0(m * std::sqrt(d)) : 1.0;
    }
  }
}

// Scalar window psi(x), psi'(x), psi''(x) of kernel/ndft-parallel.c:2288-2464
// (Kaiser-Bessel, Gaussian, B-spline, sinc-power).
inline double psi(int win, long n, double b, int m, double x) {
  const double xn = double(n) * x;
  switch (win) {
    case WIN_GAUSS: return std::exp(-(xn * xn) / b) / std::sqrt(kPi * b);
    case WIN_BSPLINE: return bspline(2 * m, xn + m);
    case WIN_SINCPOW: return std::pow(sinc_ref<double>(kPi * xn / b), 2.0 * m) / b;
    case WIN_BESSEL_I0: {  // window_bessel_i0_1d, kernel/ndft-parallel.c:2195-2203
      const double d = double(m) * m - xn * xn;
      return (d < 0) ? 0.0 : 0.5 * bessel_i0(b * std::sqrt(d));
    }
    default: return kb_psi<double>(xn, b, m);
  }
}

inline double dpsi(int win, long n, double b, int m, double x) {
  const double xn = double(n) * x;
  switch (win) {
    case WIN_GAUSS: return -2.0 * n / b * xn * std::exp(-(xn * xn) / b) / std::sqrt(kPi * b);
    case WIN_BSPLINE: return double(n) * (bspline(2 * m - 1, xn + m) - bspline(2 * m - 1, xn + m - 1));
    case WIN_SINCPOW: {
      const double y = kPi * xn / b;
      if (std::fabs(y) > Consts<double>::eps)
        return 2.0 * m * kPi * double(n) / b * (1.0 / std::tan(y) - 1.0 / y) * psi(win, n, b, m, x);
      return 0.0;
    }
    case WIN_BESSEL_I0: {  // window_bessel_i0_derivative_1d, :2205-2218
      const double d = double(m) * m - xn * xn;
      if (d < 0) return 0.0;
      const double r = std::sqrt(d);
      return (d > 0) ? -0.5 * b * double(n) * xn * bessel_i1(b * r) / r : -(0.5 * b * n) * (0.5 * b * n) * xn / n;
    }
    default: return kb_dpsi<double>(xn, double(n), b, m, kb_psi<double>(xn, b, m));
  }
}

inline double ddpsi(int win, long n, double b, int m, double x) {
  const double xn = double(n) * x;
  const double nd = double(n);
  switch (win) {
    case WIN_GAUSS: {
      const double y = xn * xn / b;
      return 2.0 * nd / b * nd * std::exp(-y) / std::sqrt(kPi * b) * (2.0 * y - 1.0);
    }
    case WIN_BSPLINE:
      return nd * nd *
             (bspline(2 * m - 2, xn + m) - 2.0 * bspline(2 * m - 2, xn + m - 1) + bspline(2 * m - 2, xn + m - 2));
    case WIN_SINCPOW: {
      const double y = kPi * xn / b;
      if (std::fabs(y) > Consts<double>::eps) {
        const double t = 1.0 / std::tan(y);
        return 2.0 * m * kPi * nd / b * (t - 1.0 / y) * dpsi(win, n, b, m, x) +
               2.0 * m * (kPi * nd / b) * (kPi * nd / b) * (-1.0 + 1.0 / (y * y) - t * t) * psi(win, n, b, m, x);
      }
      return -2.0 * m / (3.0 * b) * (kPi * nd / b) * (kPi * nd / b);
    }
    case WIN_BESSEL_I0: {  // window_bessel_i0_second_derivative_1d, :2220-2237
      const double d = double(m) * m - xn * xn;
      if (d < 0) return 0.0;
      if (d > 0) {
        const double r = std::sqrt(d), y = xn * xn;
        return 0.5 * b * nd * nd / d * (b * y * bessel_i0(b * r) - bessel_i1(b * r) / r * (y + double(m) * m));
      }
      const double q = b * b * m * nd, h = b * nd;
      return q * q / 16.0 - h * h / 4.0;
    }
    default: {
      const double p = kb_psi<double>(xn, b, m);
      const double dp = kb_dpsi<double>(xn, nd, b, m, p);
      const double d = double(m) * m - xn * xn;
      const double r = std::sqrt(std::fabs(d));
      if (d < 0)
        return 3.0 * nd * xn * dp / d + nd * nd * p / d * (1.0 + (b * xn) * (b * xn)) -
               b * nd * nd / (kPi * d) * std::cos(b * r);
      if (d > 0)
        return 3.0 * nd * xn * dp / d + nd * nd * p / d * (1.0 + (b * xn) * (b * xn)) -
               b * nd * nd / (kPi * d) * std::cosh(b * r);
      return b * (b * nd) * (b * nd) / (15.0 * kPi) * ((b * m) * (b * m) - 5.0);
    }
  }
}

}  // namespace host
}  // namespace pnfft_amd
