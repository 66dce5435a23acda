/*
 * pnfft_oracle.c -- FP64 CPU restatement of PNFFT's single-process NFFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark, and
 * never by the product library.  Every routine cites the reference lines it
 * restates (paths relative to the reference tree).
 *
 * Conventions (single process, x_max = 1/2 unless stated):
 *   f_hat  row-major [k0][k1][k2], k_t in [-N_t/2, N_t/2), index k_t + N_t/2
 *          (tests/check_vs_pfft.c:174-178, kernel/ipnfft.h:66)
 *   x      interleaved x[3j+t] in [-1/2, 1/2)             (kernel/ndft-parallel.c:2734)
 *   f      complex f[j]; grad_f complex grad[3j+t]         (kernel/ndft-parallel.c:2800-2818)
 *   grid   ghost-padded FFT output g2 of size (no_t + 2m)^3, entry for
 *          frequency-domain-shifted index l in [-no/2, no/2) at l + no/2 + m
 *          (kernel/ndft-parallel.c:1550-1572, 2545-2558)
 * Complex arrays are interleaved (re, im) doubles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
#define ORC_EPS 2.220446049250313080847e-16 /* DBL_EPSILON: kernel/ipnfft.h:99 */

/* flag values of api/pnfft.h:302-390 */
#define F_FAST_GAUSSIAN    (1u << 1)
#define F_WINDOW_GAUSSIAN  (1u << 13)
#define F_WINDOW_BSPLINE   (1u << 14)
#define F_WINDOW_SINC_POW  (1u << 15)
#define F_WINDOW_BESSEL_I0 (1u << 16)
#define F_USE_FK_GAUSS_T   (1u << 17)
#define C_COMPUTE_F        (1u << 0)
#define C_COMPUTE_GRAD_F   (1u << 1)
#define C_COMPUTE_HESSIAN  (1u << 2)

enum { W_KB = 0, W_GAUSS = 1, W_BSPLINE = 2, W_SINCPOW = 3, W_BESSEL_I0 = 4, W_GAUSS_T = 5 };

/* Window selection order of kernel/ndft-parallel.c:1650-1675. */
int orc_window_id(unsigned flags)
{
  if (flags & F_WINDOW_GAUSSIAN) return W_GAUSS;
  if (flags & F_WINDOW_BSPLINE) return W_BSPLINE;
  if (flags & F_WINDOW_SINC_POW) return W_SINCPOW;
  if (flags & F_WINDOW_BESSEL_I0) return W_BESSEL_I0;
  return W_KB;
}

/* Shape parameter b per axis: kernel/ndft-parallel.c:1009-1064 (sigma = n/N, :917-920). */
void orc_window_param_b(const long *N, const long *n, int m, unsigned flags, double *b)
{
  int w = orc_window_id(flags);
  for (int t = 0; t < 3; t++) {
    double sigma = (double)n[t] / (double)N[t];
    switch (w) {
      case W_GAUSS:   b[t] = ((double)m / ORC_PI) * 2.0 * sigma / (2.0 * sigma - 1.0); break;
      case W_BSPLINE: b[t] = 0.0; break;
      case W_SINCPOW: b[t] = (double)m * (2.0 * sigma) / (2.0 * sigma - 1.0); break;
      default:        b[t] = ORC_PI * (2.0 - 1.0 / sigma); break; /* KB and Bessel-I0 */
    }
  }
}

/* ---------------------------------------------------------------------------
 * Special functions.
 * ------------------------------------------------------------------------- */

/* Modified Bessel function I0 in FP64.  The reference forwards to GSL's
 * gsl_sf_bessel_I0 (kernel/bessel_i0.c:372-375).  Restated from the
 * mathematics: the power series sum_k (x^2/4)^k/(k!)^2 (all terms positive, so
 * no cancellation) for |x| <= 40, and the Hankel asymptotic expansion
 * e^x/sqrt(2 pi x) * sum_k ((2k-1)!!)^2/(k! 8^k x^k) beyond (Abramowitz &
 * Stegun 9.7.1), truncated at its smallest term. */
double orc_bessel_i0(double x)
{
  double ax = fabs(x);
  if (ax <= 40.0) {
    double q = 0.25 * ax * ax, term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; k++) {
      term *= q / ((double)k * (double)k);
      sum += term;
      if (term < 1e-18 * sum) break;
    }
    return sum;
  }
  double term = 1.0, sum = 1.0, prev = 1.0;
  for (int k = 1; k < 60; k++) {
    double c = (2.0 * k - 1.0);
    term *= c * c / (8.0 * k * ax);
    if (term > prev) break;
    sum += term;
    prev = term;
    if (term < 1e-18 * sum) break;
  }
  return exp(ax) / sqrt(2.0 * ORC_PI * ax) * sum;
}

/* Modified Bessel function I1 in FP64 (the reference calls GSL,
 * kernel/bessel_i1.c:567).  Restated from the mathematics: the power series
 * (x/2) sum_k (x^2/4)^k/(k!(k+1)!) for |x| <= 40, and the Hankel expansion
 * e^x/sqrt(2 pi x) sum_k t_k, t_k = t_{k-1} ((2k-1)^2 - 4)/(8 k x) beyond
 * (Abramowitz & Stegun 9.7.1), truncated at its smallest term.  Odd in x. */
double orc_bessel_i1(double x)
{
  double ax = fabs(x), r;
  if (ax <= 40.0) {
    double q = 0.25 * ax * ax, term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; k++) {
      term *= q / ((double)k * (double)(k + 1));
      sum += term;
      if (term < 1e-18 * sum) break;
    }
    r = 0.5 * ax * sum;
  } else {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 60; k++) {
      double c = (2.0 * k - 1.0);
      double next = term * (c * c - 4.0) / (8.0 * k * ax);
      if (fabs(next) >= fabs(term)) break;
      term = next;
      sum += term;
      if (fabs(term) < 1e-18 * fabs(sum)) break;
    }
    r = exp(ax) / sqrt(2.0 * ORC_PI * ax) * sum;
  }
  return x < 0 ? -r : r;
}

/* sin(x)/x with the Taylor branches of kernel/sinc.c:26-52 (thresholds
 * eps, eps^(1/2), eps^(1/4) with eps = DBL_EPSILON). */
double orc_sinc(double x)
{
  const double b = ORC_EPS, bs = sqrt(b), bs2 = sqrt(bs);
  if (fabs(x) >= bs2) return sin(x) / x;
  double r = 1.0;
  if (fabs(x) >= b) {
    double x2 = x * x;
    r -= x2 / 6.0;
    if (fabs(x) >= bs) r += (x2 * x2) / 120.0;
  }
  return r;
}

/* All nonzero values V[d] = M_k(u + d), d = 0..k-1, of the cardinal B-spline
 * M_k (support [0,k]) for u in (0,1], by the recursion
 *   M_r(t) = (t M_{r-1}(t) + (r - t) M_{r-1}(t - 1)) / (r - 1).
 * V must hold k doubles. */
static void bspline_column(int k, double u, double *V)
{
  V[0] = 1.0;
  for (int r = 2; r <= k; r++) {
    double inv = 1.0 / (double)(r - 1);
    V[r - 1] = 0.0;
    for (int d = r - 1; d >= 0; d--) {
      double left = (d > 0) ? V[d - 1] : 0.0;
      V[d] = ((u + d) * V[d] + ((double)r - u - d) * left) * inv;
    }
  }
}

/* Cardinal B-spline M_k(x), zero outside (0, k).  Same function as the de Boor
 * scheme of kernel/bspline.c:45-91 (PNX(bspline)). */
double orc_bspline(int k, double x)
{
  if (!(x > 0.0 && x < (double)k)) return 0.0;
  double V[64];
  if (k > 64) return NAN;
  double fl = floor(x);
  double u = x - fl;
  int d0 = (int)fl;
  if (u == 0.0) { u = 1.0; d0 -= 1; }
  bspline_column(k, u, V);
  return V[d0];
}

/* Complex error function erf(x + i y), x > 0, by Abramowitz & Stegun 7.1.29:
 *   erf(x) + e^{-x^2}/(2 pi x) [(1 - cos 2xy) + i sin 2xy]
 *   + (2/pi) e^{-x^2} sum_{k>=1} e^{-k^2/4}/(k^2 + 4x^2) [f_k + i g_k],
 *   f_k = 2x - 2x cosh(ky) cos(2xy) + k sinh(ky) sin(2xy),
 *   g_k = 2x cosh(ky) sin(2xy) + k sinh(ky) cos(2xy),
 * relative error about 1e-16 (the reference calls libcerf's cerf, cerf/). */
void orc_cerf(double x, double y, double *re, double *im)
{
  const double e = exp(-x * x), c = cos(2.0 * x * y), sn = sin(2.0 * x * y);
  double r = erf(x) + e / (2.0 * ORC_PI * x) * (1.0 - c);
  double i = e / (2.0 * ORC_PI * x) * sn;
  double sr = 0.0, si = 0.0;
  for (int k = 1; k <= 60; k++) {
    const double w = exp(-0.25 * k * k) / (k * k + 4.0 * x * x);
    const double ch = cosh(k * y), sh = sinh(k * y);
    sr += w * (2.0 * x - 2.0 * x * ch * c + k * sh * sn);
    si += w * (2.0 * x * ch * sn + k * sh * c);
  }
  *re = r + 2.0 / ORC_PI * e * sr;
  *im = i + 2.0 / ORC_PI * e * si;
}

/* ---------------------------------------------------------------------------
 * Window Fourier coefficients (matrix D), kernel/matrix_D.c:30-225.
 * The factor 1/n of D cancels the factor n of the coefficients (:29).
 * ------------------------------------------------------------------------- */
double orc_inv_phi_hat(int window, long k, long n, double b, int m)
{
  switch (window) {
    case W_GAUSS: { /* PNFFT_INV_PHI_HAT_GAUSS, :30-31 */
      double a = ORC_PI * (double)k / (double)n;
      return exp(a * a * b);
    }
    case W_BSPLINE: /* PNFFT_INV_PHI_HAT_BSPLINE, :118-119 */
      return pow(orc_sinc((double)k * ORC_PI / (double)n), -2.0 * m);
    case W_SINCPOW: { /* inv_phi_hat_sinc_power, :124-132 */
      double d = fabs((double)k * b / (double)n);
      return (d < m) ? 1.0 / orc_bspline(2 * m, d + m) : 0.0;
    }
    case W_GAUSS_T: { /* inv_phi_hat_gauss_t, :37-55: 1 / (Re erf(m/sqrt b + i pi k sqrt b / n) e^{-b (pi k/n)^2}) */
      double sb = sqrt(b), re, im;
      orc_cerf((double)m / sb, ORC_PI * (double)k * sb / (double)n, &re, &im);
      double a = ORC_PI * (double)k / (double)n;
      return 1.0 / (re * exp(-a * a * b));
    }
    case W_BESSEL_I0: { /* inv_phi_hat_bessel_i0, :91-102 */
      double a = 2.0 * ORC_PI * (double)k / (double)n;
      double d = b * b - a * a;
      double r = sqrt(fabs(d));
      if (d < 0) return r / sin(m * r);
      return (d > 0) ? r / sinh(m * r) : 1.0 / m;
    }
    default: { /* inv_phi_hat_kaiser, :66-77 */
      double a = 2.0 * ORC_PI * (double)k / (double)n;
      double d = b * b - a * a;
      if (d < 0) return 0.0;
      return (d > 0) ? 1.0 / orc_bessel_i0(m * sqrt(d)) : 1.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Window values on the (2m+1)-point stencil of one axis.
 *   psi[s]  = psi(x - l_s/n), l_s = floor(n x) - m + s, s = 0..2m
 *   dpsi[s] = psi'(x - l_s/n)   (derivative with respect to x)
 * Restates pre_psi_tensor_* (kernel/ndft-parallel.c:1678-1797) and
 * pre_dpsi_tensor_* (:1850-1953) with the 1-d kernels of :2241-2269.
 * nx is n*x computed in FP64, fl = floor(nx).
 * ------------------------------------------------------------------------- */
static double kb_1d(double xn, double b, int m) /* kaiser_bessel_1d, :2241-2254; xn = x*n */
{
  double d = (double)m * m - xn * xn;
  double r = sqrt(fabs(d));
  if (d < 0) return sin(b * r) / (ORC_PI * r);
  return (d > 0) ? sinh(b * r) / (ORC_PI * r) : b / ORC_PI;
}

static double kb_d1d(double xn, long n, double b, int m, double psi) /* :2256-2269 */
{
  double d = (double)m * m - xn * xn;
  double r = sqrt(fabs(d));
  /* n*n*x/d with x = xn/n */
  if (d < 0) return (double)n * xn / d * (psi - b * cos(b * r) / ORC_PI);
  return (d > 0) ? (double)n * xn / d * (psi - b * cosh(b * r) / ORC_PI)
                 : -(double)n * m * b * b * b / (3.0 * ORC_PI);
}

/* Bessel-I0 window, window_bessel_i0_1d / _derivative_1d of
 * kernel/ndft-parallel.c:2195-2218 with xn = x*n: I0(b sqrt(m^2 - xn^2))/2 on
 * the support, -(b/2) n xn I1(b r)/r for the derivative (-(b n/2)^2 xn/n at
 * r = 0). */
static double bi0_1d(double xn, double b, int m)
{
  double d = (double)m * m - xn * xn;
  return (d < 0) ? 0.0 : 0.5 * orc_bessel_i0(b * sqrt(d));
}

static double bi0_d1d(double xn, long n, double b, int m)
{
  double d = (double)m * m - xn * xn;
  if (d < 0) return 0.0;
  double r = sqrt(d);
  return (d > 0) ? -0.5 * b * (double)n * xn * orc_bessel_i1(b * r) / r
                 : -(0.5 * b * n) * (0.5 * b * n) * xn / (double)n;
}

/* Second derivatives on the support (kernel/ndft-parallel.c:2220-2237 for the
 * Bessel-I0 window, :2272-2285 for Kaiser-Bessel), xn = x*n. */
static double bi0_dd1d(double xn, long n, double b, int m)
{
  double d = (double)m * m - xn * xn;
  if (d < 0) return 0.0;
  if (d > 0) {
    double r = sqrt(d), y = xn * xn;
    return 0.5 * b * (double)n * n / d * (b * y * orc_bessel_i0(b * r) - orc_bessel_i1(b * r) / r * (y + (double)m * m));
  }
  double q = b * b * m * (double)n, h = b * (double)n;
  return q * q / 16.0 - h * h / 4.0;
}

static double kb_dd1d(double xn, long n, double b, int m, double psi, double dpsi)
{
  double d = (double)m * m - xn * xn;
  double r = sqrt(fabs(d));
  double nn = (double)n * n, bx = b * xn;
  if (d == 0) {
    double bn = b * n, bm = b * m;
    return b * bn * bn / (15.0 * ORC_PI) * (bm * bm - 5.0);
  }
  double hyp = (d < 0) ? cos(b * r) : cosh(b * r);
  return 3.0 * n * xn * dpsi / d + nn * psi / d * (1.0 + bx * bx) - b * nn / (ORC_PI * d) * hyp;
}

/* psi'' on the stencil, pre_ddpsi_tensor_* of kernel/ndft-parallel.c:1956-2105
 * (psi and dpsi of the same node as inputs, as the reference passes them). */
static void ddpsi_axis(int window, long n, double b, int m, double nx, const double *psi,
                       const double *dpsi, double *ddpsi)
{
  const int c = 2 * m + 1;
  const double fl = floor(nx);
  const double uj = nx - fl + m;
  const double nn = (double)n * n;
  switch (window) {
    case W_GAUSS: /* :2010-2024 */
      for (int s = 0; s < c; s++)
        ddpsi[s] = 2.0 * nn / b * (2.0 / b * (uj - s) * (uj - s) - 1.0) * psi[s];
      break;
    case W_BSPLINE: { /* :2026-2049: n^2 (M_{2m-2}(t) - 2 M_{2m-2}(t-1) + M_{2m-2}(t-2)) */
      double Q[64];
      bspline_column(2 * m - 2, 1.0 - (nx - fl), Q); /* Q[d] = M_{2m-2}(1 - frac + d) */
      for (int s = 0; s < c; s++) {
        double q3 = (s >= 3 && s - 3 <= 2 * m - 3) ? Q[s - 3] : 0.0;
        double q2 = (s >= 2 && s - 2 <= 2 * m - 3) ? Q[s - 2] : 0.0;
        double q1 = (s >= 1 && s - 1 <= 2 * m - 3) ? Q[s - 1] : 0.0;
        ddpsi[s] = nn * (q3 - 2.0 * q2 + q1);
      }
      break;
    }
    case W_SINCPOW: /* :2052-2071 */
      for (int s = 0; s < c; s++) {
        double y = ORC_PI * (uj - s) / b;
        double a = ORC_PI * (double)n / b;
        if (fabs(y) > ORC_EPS) {
          double t = tan(y);
          ddpsi[s] = 2.0 * m * a * (1.0 / t - 1.0 / y) * dpsi[s] +
                     2.0 * m * a * a * (1.0 / (y * y) - 1.0 - 1.0 / (t * t)) * psi[s];
        } else {
          ddpsi[s] = -2.0 * m * a * a / (3.0 * b);
        }
      }
      break;
    case W_BESSEL_I0: /* :2073-2089 */
      for (int s = 0; s < c; s++) ddpsi[s] = bi0_dd1d(uj - s, n, b, m);
      break;
    default: /* Kaiser-Bessel, :2091-2105 */
      for (int s = 0; s < c; s++) ddpsi[s] = kb_dd1d(uj - s, n, b, m, psi[s], dpsi[s]);
      break;
  }
}

void orc_psi_axis(int window, long n, double b, int m, double nx,
                  double *psi, double *dpsi)
{
  const int c = 2 * m + 1;
  const double fl = floor(nx);
  const double frac = nx - fl;
  switch (window) {
    case W_GAUSS: {
      double norm = sqrt(ORC_PI * b);
      for (int s = 0; s < c; s++) {
        double u = (fl - nx - m) + s; /* u_j + s, :1688-1690 */
        psi[s] = exp(-(u * u) / b) / norm;
      }
      if (dpsi) {
        double uj = nx - fl + m; /* :1860-1862 */
        for (int s = 0; s < c; s++) dpsi[s] = -2.0 * n / b * (uj - s) * psi[s];
      }
      break;
    }
    case W_BSPLINE: {
      /* psi[s] = M_{2m}(s - frac) (:1724-1741), dpsi[s] = -n M_{2m}'(s - frac)
       * (:1874-1892); psi[0] is always zero (fast_bspline(-1,..) = 0). */
      double V[64], W[64];
      double u = 1.0 - frac;
      bspline_column(2 * m, u, V);
      bspline_column(2 * m - 1, u, W);
      psi[0] = 0.0;
      for (int s = 1; s < c; s++) psi[s] = V[s - 1];
      if (dpsi) {
        for (int s = 0; s < c; s++) {
          double a = (s >= 1 && s - 1 < 2 * m - 1) ? W[s - 1] : 0.0;
          double bb = (s >= 2 && s - 2 < 2 * m - 1) ? W[s - 2] : 0.0;
          dpsi[s] = -(double)n * (a - bb);
        }
      }
      break;
    }
    case W_SINCPOW: {
      for (int s = 0; s < c; s++) {
        double u = (fl - nx - m) + s; /* :1755-1761 */
        psi[s] = pow(orc_sinc(ORC_PI * u / b), 2.0 * m) / b;
      }
      if (dpsi) {
        double uj = nx - fl + m; /* :1907-1915 */
        for (int s = 0; s < c; s++) {
          double y = ORC_PI * (uj - s) / b;
          dpsi[s] = (fabs(y) > ORC_EPS)
                        ? 2.0 * m * ORC_PI * (double)n / b * (1.0 / tan(y) - 1.0 / y) * psi[s]
                        : 0.0;
        }
      }
      break;
    }
    case W_BESSEL_I0: { /* pre_psi_tensor_bessel_i0 :1765-1781, pre_dpsi :1919-1935 */
      double uj = nx - fl + m;
      for (int s = 0; s < c; s++) psi[s] = bi0_1d(uj - s, b, m);
      if (dpsi)
        for (int s = 0; s < c; s++) dpsi[s] = bi0_d1d(uj - s, n, b, m);
      break;
    }
    default: { /* Kaiser-Bessel, :1791-1796 and :1946-1951 */
      for (int s = 0; s < c; s++) {
        double u = (fl - nx - m) + s;
        psi[s] = kb_1d(u, b, m);
      }
      if (dpsi) {
        double uj = nx - fl + m;
        for (int s = 0; s < c; s++) dpsi[s] = kb_d1d(uj - s, n, b, m, psi[s]);
      }
      break;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Interpolated windows, PNFFT_PRE_{CONST,LIN,QUAD,CUB}_PSI
 * (kernel/ndft-parallel.c:1088-1122, :318-352, :1586-1617, kernel/ipnfft.h:457-497).
 * A table of axis t holds, for k in [0, K), stencil point c and offsets
 * i = -order/2 .. (order+1)/2, the value psi^(d)((m + (k+i)/K - c)/n);
 * a node with fractional part dist = n x - floor(n x) reads the entries of
 * k = floor(dist K) and interpolates at dist K - k.  orc_intpol_set makes
 * node_windows (and so every B routine) read tables instead of evaluating.
 * ------------------------------------------------------------------------- */
double orc_ddpsi(int window, long n, double b, int m, double x);
double orc_dpsi(int window, long n, double b, int m, double x);
double orc_psi(int window, long n, double b, int m, double x);

void orc_intpol_table(int window, long n, double b, int m, int order, long K, int deriv, double *table)
{
  const int cutoff = 2 * m + 1;
  long ind = 0;
  for (long k = 0; k < K; k++)
    for (int c = 0; c < cutoff; c++)
      for (int i = -order / 2; i <= (order + 1) / 2; i++, ind++) {
        const double x = ((double)m + (double)(k + i) / (double)K - c) / (double)n;
        table[ind] = deriv == 0 ? orc_psi(window, n, b, m, x)
                   : deriv == 1 ? orc_dpsi(window, n, b, m, x) : orc_ddpsi(window, n, b, m, x);
      }
}

static double intpol_eval(int order, const double *f, double dist)
{
  switch (order) {
    case 0: return f[0];
    case 1: return f[0] * (1.0 - dist) + f[1] * dist;
    case 2: {
      const double c0 = dist + 1.0, c1 = dist, c2 = dist - 1.0;
      return 0.5 * f[0] * c1 * c2 - c0 * f[1] * c2 + 0.5 * c0 * c1 * f[2];
    }
    default: {
      const double c0 = dist + 1.0, c1 = dist, c2 = dist - 1.0, c3 = dist - 2.0;
      return (-f[0] * c1 * c2 * c3 + 3.0 * c0 * f[1] * c2 * c3 - 3.0 * c0 * c1 * f[2] * c3 + c0 * c1 * c2 * f[3]) / 6.0;
    }
  }
}

static struct {
  int order;
  long K;
  const double *tab[3][3]; /* [derivative][axis] */
} g_ip = {-1, 0, {{0}}};

void orc_intpol_set(int order, long K, const double *p0, const double *p1, const double *p2, const double *d0,
                    const double *d1, const double *d2, const double *dd0, const double *dd1, const double *dd2)
{
  g_ip.order = order;
  g_ip.K = K;
  const double *t[9] = {p0, p1, p2, d0, d1, d2, dd0, dd1, dd2};
  for (int q = 0; q < 9; q++) g_ip.tab[q / 3][q % 3] = t[q];
}

void orc_intpol_clear(void) { g_ip.order = -1; }

/* pre_tensor_intpol (kernel/ndft-parallel.c:1586-1617) for axis t, derivative d */
void orc_intpol_axis(int t, int d, int m, double nx, double *out)
{
  const int cutoff = 2 * m + 1, w = g_ip.order + 1;
  const double dist = nx - floor(nx);
  const long k = (long)floor(dist * (double)g_ip.K);
  const double dk = dist * (double)g_ip.K - (double)k;
  const double *tab = g_ip.tab[d][t] + k * cutoff * w;
  for (int s = 0; s < cutoff; s++) out[s] = intpol_eval(g_ip.order, tab + s * w, dk);
}

/* psi, psi' and psi'' of one axis's stencil (ddpsi may not be null). */
void orc_psi_axis2(int window, long n, double b, int m, double nx, double *psi, double *dpsi,
                   double *ddpsi)
{
  orc_psi_axis(window, n, b, m, nx, psi, dpsi);
  ddpsi_axis(window, n, b, m, nx, psi, dpsi, ddpsi);
}

/* Scalar window psi(x) and psi'(x) (PNX(psi), PNX(dpsi): :2288-2318). */
double orc_psi(int window, long n, double b, int m, double x)
{
  double xn = (double)n * x;
  switch (window) {
    case W_GAUSS: return exp(-(xn * xn) / b) / sqrt(ORC_PI * b);
    case W_BSPLINE: return orc_bspline(2 * m, xn + m);
    case W_SINCPOW: return pow(orc_sinc(ORC_PI * xn / b), 2.0 * m) / b;
    case W_BESSEL_I0: return bi0_1d(xn, b, m);
    default: return kb_1d(xn, b, m);
  }
}

double orc_dpsi(int window, long n, double b, int m, double x);

/* Scalar psi''(x) (PNX(ddpsi), kernel/ndft-parallel.c:2320-2340, 2350-2460). */
double orc_ddpsi(int window, long n, double b, int m, double x)
{
  double xn = (double)n * x;
  double nn = (double)n * n;
  switch (window) {
    case W_GAUSS: {
      double y = xn * xn / b;
      return 2.0 * n / b * n * exp(-y) / sqrt(ORC_PI * b) * (2.0 * y - 1.0);
    }
    case W_BSPLINE:
      return nn * (orc_bspline(2 * m - 2, xn + m) - 2.0 * orc_bspline(2 * m - 2, xn + m - 1) +
                   orc_bspline(2 * m - 2, xn + m - 2));
    case W_SINCPOW: {
      double y = ORC_PI * xn / b, a = ORC_PI * (double)n / b;
      if (fabs(y) > ORC_EPS) {
        double t = tan(y);
        return 2.0 * m * a * (1.0 / t - 1.0 / y) * orc_dpsi(window, n, b, m, x) +
               2.0 * m * a * a * (1.0 / (y * y) - 1.0 - 1.0 / (t * t)) * orc_psi(window, n, b, m, x);
      }
      return -2.0 * m / (3.0 * b) * a * a;
    }
    case W_BESSEL_I0: return bi0_dd1d(xn, n, b, m);
    default: {
      double p = kb_1d(xn, b, m);
      return kb_dd1d(xn, n, b, m, p, kb_d1d(xn, n, b, m, p));
    }
  }
}

double orc_dpsi(int window, long n, double b, int m, double x)
{
  double xn = (double)n * x;
  switch (window) {
    case W_GAUSS: return -2.0 * n / b * xn * exp(-(xn * xn) / b) / sqrt(ORC_PI * b);
    case W_BSPLINE:
      return (double)n * (orc_bspline(2 * m - 1, xn + m) - orc_bspline(2 * m - 1, xn + m - 1));
    case W_SINCPOW: {
      double y = ORC_PI * xn / b;
      if (fabs(y) > ORC_EPS)
        return 2.0 * m * ORC_PI * (double)n / b * (1.0 / tan(y) - 1.0 / y) *
               orc_psi(window, n, b, m, x);
      return 0.0;
    }
    case W_BESSEL_I0: return bi0_d1d(xn, n, b, m);
    default: return kb_d1d(xn, n, b, m, kb_1d(xn, b, m));
  }
}

/* ---------------------------------------------------------------------------
 * NDFT (direct sums), kernel/ndft-parallel.c:377-722, c2c.
 * The reference forms e^{-2 pi i k x} by incremental products (:443-449,
 * :456-491); here each axis factor is taken from cos/sin directly and the
 * triple sum is nested per axis, which is the same sum with smaller rounding.
 * trafo overwrites f / grad (the NDFT zeroes its outputs, :400-414); adj adds
 * onto f_hat (:649-650).
 * ------------------------------------------------------------------------- */
void orc_ndft_trafo_c2c(const long *N, long M, const double *x, const double *fhat,
                        double *f, double *grad, unsigned cflags)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2];
#pragma omp parallel
  {
    double *e0 = malloc(sizeof(double) * 2 * N0);
    double *e1 = malloc(sizeof(double) * 2 * N1);
    double *e2 = malloc(sizeof(double) * 2 * N2);
#pragma omp for schedule(dynamic, 16)
    for (long j = 0; j < M; j++) {
      const long Nt[3] = {N0, N1, N2};
      double *et[3] = {e0, e1, e2};
      for (int t = 0; t < 3; t++)
        for (long i = 0; i < Nt[t]; i++) {
          double k = (double)(i - Nt[t] / 2);
          double ph = -2.0 * ORC_PI * k * x[3 * j + t];
          et[t][2 * i] = cos(ph);
          et[t][2 * i + 1] = sin(ph);
        }
      double fr = 0, fi = 0, g[6] = {0, 0, 0, 0, 0, 0};
      for (long i0 = 0; i0 < N0; i0++) {
        double s1r = 0, s1i = 0, s1kr = 0, s1ki = 0, s1zr = 0, s1zi = 0;
        for (long i1 = 0; i1 < N1; i1++) {
          const double *fh = fhat + 2 * ((i0 * N1 + i1) * N2);
          double s2r = 0, s2i = 0, s2kr = 0, s2ki = 0;
          for (long i2 = 0; i2 < N2; i2++) {
            double ar = fh[2 * i2], ai = fh[2 * i2 + 1];
            double br = e2[2 * i2], bi = e2[2 * i2 + 1];
            double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
            double k2 = (double)(i2 - N2 / 2);
            s2r += pr; s2i += pi;
            s2kr += k2 * pr; s2ki += k2 * pi;
          }
          double br = e1[2 * i1], bi = e1[2 * i1 + 1];
          double k1 = (double)(i1 - N1 / 2);
          double tr = s2r * br - s2i * bi, ti = s2r * bi + s2i * br;
          s1r += tr; s1i += ti;
          s1kr += k1 * tr; s1ki += k1 * ti;
          s1zr += s2kr * br - s2ki * bi; s1zi += s2kr * bi + s2ki * br;
        }
        double br = e0[2 * i0], bi = e0[2 * i0 + 1];
        double k0 = (double)(i0 - N0 / 2);
        double tr = s1r * br - s1i * bi, ti = s1r * bi + s1i * br;
        fr += tr; fi += ti;
        g[0] += k0 * tr; g[1] += k0 * ti;
        g[2] += s1kr * br - s1ki * bi; g[3] += s1kr * bi + s1ki * br;
        g[4] += s1zr * br - s1zi * bi; g[5] += s1zr * bi + s1zi * br;
      }
      if ((cflags & C_COMPUTE_F) && f) { f[2 * j] = fr; f[2 * j + 1] = fi; }
      if ((cflags & C_COMPUTE_GRAD_F) && grad) {
        /* grad *= -2 pi i (:594-603) */
        for (int t = 0; t < 3; t++) {
          double re = g[2 * t], im = g[2 * t + 1];
          grad[6 * j + 2 * t] = 2.0 * ORC_PI * im;
          grad[6 * j + 2 * t + 1] = -2.0 * ORC_PI * re;
        }
      }
    }
    free(e0); free(e1); free(e2);
  }
}

void orc_ndft_adj_c2c(const long *N, long M, const double *x, const double *f,
                      const double *grad, double *fhat, unsigned cflags)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2];
  const long Ntot = N0 * N1 * N2;
  /* Parallel over k0 slabs: each thread owns f_hat[k0][*][*]. */
#pragma omp parallel
  {
    double *e1 = malloc(sizeof(double) * 2 * N1);
    double *e2 = malloc(sizeof(double) * 2 * N2);
#pragma omp for schedule(dynamic, 1)
    for (long i0 = 0; i0 < N0; i0++) {
      double k0 = (double)(i0 - N0 / 2);
      for (long j = 0; j < M; j++) {
        double wr = 0, wi = 0; /* weight multiplying e^{+2 pi i k x} */
        double gk[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        int use_f = (cflags & C_COMPUTE_F) && f;
        int use_g = (cflags & C_COMPUTE_GRAD_F) && grad;
        if (use_f) { wr = f[2 * j]; wi = f[2 * j + 1]; }
        if (use_g)
          for (int t = 0; t < 3; t++) { gk[t][0] = grad[6 * j + 2 * t]; gk[t][1] = grad[6 * j + 2 * t + 1]; }
        double ph0 = 2.0 * ORC_PI * k0 * x[3 * j];
        double c0 = cos(ph0), s0 = sin(ph0);
        for (long i1 = 0; i1 < N1; i1++) {
          double ph = 2.0 * ORC_PI * (double)(i1 - N1 / 2) * x[3 * j + 1];
          e1[2 * i1] = cos(ph); e1[2 * i1 + 1] = sin(ph);
        }
        for (long i2 = 0; i2 < N2; i2++) {
          double ph = 2.0 * ORC_PI * (double)(i2 - N2 / 2) * x[3 * j + 2];
          e2[2 * i2] = cos(ph); e2[2 * i2 + 1] = sin(ph);
        }
        for (long i1 = 0; i1 < N1; i1++) {
          double k1 = (double)(i1 - N1 / 2);
          double e01r = c0 * e1[2 * i1] - s0 * e1[2 * i1 + 1];
          double e01i = c0 * e1[2 * i1 + 1] + s0 * e1[2 * i1];
          double *fh = fhat + 2 * ((i0 * N1 + i1) * N2);
          for (long i2 = 0; i2 < N2; i2++) {
            double k2 = (double)(i2 - N2 / 2);
            double er = e01r * e2[2 * i2] - e01i * e2[2 * i2 + 1];
            double ei = e01r * e2[2 * i2 + 1] + e01i * e2[2 * i2];
            /* coefficient: f_j + 2 pi i (k . g_j)   (:681, :699-707) */
            double cr = wr, ci = wi;
            if (use_g) {
              double sr = k0 * gk[0][0] + k1 * gk[1][0] + k2 * gk[2][0];
              double si = k0 * gk[0][1] + k1 * gk[1][1] + k2 * gk[2][1];
              cr += -2.0 * ORC_PI * si;
              ci += 2.0 * ORC_PI * sr;
            }
            fh[2 * i2] += cr * er - ci * ei;
            fh[2 * i2 + 1] += cr * ei + ci * er;
          }
        }
      }
    }
    free(e1); free(e2);
  }
  (void)Ntot;
}

/* NDFT Hessian, c2c (kernel/ndft-parallel.c:450-510, scaled by -4 pi^2 at
 * :604-612): hess[6j + c] = -4 pi^2 sum_k k_a k_b f_hat_k e^{-2 pi i k x_j}
 * for (a, b) = xx, xy, xz, yy, yz, zz; overwrites hess. */
void orc_ndft_hessian_c2c(const long *N, long M, const double *x, const double *fhat, double *hess)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2];
  static const int A[6] = {0, 0, 0, 1, 1, 2}, B[6] = {0, 1, 2, 1, 2, 2};
#pragma omp parallel for schedule(dynamic, 16)
  for (long j = 0; j < M; j++) {
    double h[12] = {0};
    for (long i0 = 0; i0 < N0; i0++)
      for (long i1 = 0; i1 < N1; i1++)
        for (long i2 = 0; i2 < N2; i2++) {
          double k[3] = {(double)(i0 - N0 / 2), (double)(i1 - N1 / 2), (double)(i2 - N2 / 2)};
          const double *fh = fhat + 2 * ((i0 * N1 + i1) * N2 + i2);
          double ph = -2.0 * ORC_PI * (k[0] * x[3 * j] + k[1] * x[3 * j + 1] + k[2] * x[3 * j + 2]);
          double er = cos(ph), ei = sin(ph);
          double pr = fh[0] * er - fh[1] * ei, pi = fh[0] * ei + fh[1] * er;
          for (int q = 0; q < 6; q++) {
            double kk = k[A[q]] * k[B[q]];
            h[2 * q] += kk * pr;
            h[2 * q + 1] += kk * pi;
          }
        }
    for (int q = 0; q < 12; q++) hess[12 * j + q] = -4.0 * ORC_PI * ORC_PI * h[q];
  }
}

/* ---------------------------------------------------------------------------
 * c2r NDFT (kernel/ndft-parallel.c:356-374, :524-533, :569-573, :594-603).
 * f_hat holds the half spectrum [k0][k1][k2], k2 in [-N2/2, 0] (N2/2+1 entries).
 * ------------------------------------------------------------------------- */
static int is_hermitian(long k0, long k1, long k2, long N0, long N1, long N2)
{
  if ((k0 == 0 || k0 == -N0 / 2) && (k1 == 0 || k1 == -N1 / 2) && (k2 == 0 || k2 == -N2 / 2))
    return 1;
  if (k1 > 0 && (k2 == 0 || k2 == -N2 / 2)) return 1;
  if (k0 > 0 && (k1 == 0 || k1 == -N1 / 2) && (k2 == 0 || k2 == -N2 / 2)) return 1;
  return 0;
}

void orc_ndft_trafo_c2r(const long *N, long M, const double *x, const double *fhat,
                        double *f, double *grad, unsigned cflags)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2], H2 = N2 / 2 + 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (long j = 0; j < M; j++) {
    double fr = 0, g[3] = {0, 0, 0};
    for (long i0 = 0; i0 < N0; i0++) {
      long k0 = i0 - N0 / 2;
      for (long i1 = 0; i1 < N1; i1++) {
        long k1 = i1 - N1 / 2;
        for (long i2 = 0; i2 < H2; i2++) {
          long k2 = i2 - N2 / 2;
          const double *fh = fhat + 2 * ((i0 * N1 + i1) * H2 + i2);
          double ph = -2.0 * ORC_PI * ((double)k0 * x[3 * j] + (double)k1 * x[3 * j + 1] + (double)k2 * x[3 * j + 2]);
          double er = cos(ph), ei = sin(ph);
          double pr = fh[0] * er - fh[1] * ei, pi = fh[0] * ei + fh[1] * er;
          if (k0 == 0 && k1 == 0 && k2 == 0) {
            fr += fh[0];
          } else if (!is_hermitian(k0, k1, k2, N0, N1, N2)) {
            fr += 2.0 * pr;
            g[0] += 2.0 * k0 * pi; g[1] += 2.0 * k1 * pi; g[2] += 2.0 * k2 * pi;
          }
        }
      }
    }
    if ((cflags & C_COMPUTE_F) && f) f[j] = fr;
    if ((cflags & C_COMPUTE_GRAD_F) && grad)
      for (int t = 0; t < 3; t++) grad[3 * j + t] = -2.0 * ORC_PI * g[t];
  }
}

/* c2r NDFT Hessian (kernel/ndft-parallel.c:465-476, :494-499, :604-609):
 * -4 pi^2 sum over the half spectrum of 2 k_a k_b Re(f_hat_k e^{-2 pi i k x}),
 * skipping k = 0 and the entries is_hermitian() marks; overwrites hess. */
void orc_ndft_hessian_c2r(const long *N, long M, const double *x, const double *fhat, double *hess)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2], H2 = N2 / 2 + 1;
  static const int A[6] = {0, 0, 0, 1, 1, 2}, B[6] = {0, 1, 2, 1, 2, 2};
#pragma omp parallel for schedule(dynamic, 16)
  for (long j = 0; j < M; j++) {
    double h[6] = {0};
    for (long i0 = 0; i0 < N0; i0++)
      for (long i1 = 0; i1 < N1; i1++)
        for (long i2 = 0; i2 < H2; i2++) {
          long k0 = i0 - N0 / 2, k1 = i1 - N1 / 2, k2 = i2 - N2 / 2;
          if ((k0 == 0 && k1 == 0 && k2 == 0) || is_hermitian(k0, k1, k2, N0, N1, N2)) continue;
          double k[3] = {(double)k0, (double)k1, (double)k2};
          const double *fh = fhat + 2 * ((i0 * N1 + i1) * H2 + i2);
          double ph = -2.0 * ORC_PI * (k[0] * x[3 * j] + k[1] * x[3 * j + 1] + k[2] * x[3 * j + 2]);
          double pr = fh[0] * cos(ph) - fh[1] * sin(ph);
          for (int q = 0; q < 6; q++) h[q] += 2.0 * k[A[q]] * k[B[q]] * pr;
        }
    for (int q = 0; q < 6; q++) hess[6 * j + q] = -4.0 * ORC_PI * ORC_PI * h[q];
  }
}

/* ---------------------------------------------------------------------------
 * Deconvolution D / D^T (kernel/matrix_D.c:227-280, :319-395): separable
 * factors inv_phi_hat_t(k_t), trafo overwrites g1, adjoint accumulates f_hat.
 * tables: concatenated per-axis tables of length N0+N1+N2 (k_t ascending).
 * ------------------------------------------------------------------------- */
void orc_inv_phi_hat_tables(const long *N, const long *n, int m, unsigned flags, double *tab)
{
  double b[3];
  int w = orc_window_id(flags);
  /* the truncated Gaussian changes only D (PNX(inv_phi_hat), kernel/matrix_D.c:191-196) */
  if (w == W_GAUSS && (flags & F_USE_FK_GAUSS_T)) w = W_GAUSS_T;
  orc_window_param_b(N, n, m, flags, b);
  long off = 0;
  for (int t = 0; t < 3; t++)
    for (long i = 0; i < N[t]; i++) tab[off++] = orc_inv_phi_hat(w, i - N[t] / 2, n[t], b[t], m);
}

void orc_D(const long *N, const double *tab, const double *in, double *out, int accumulate)
{
  const double *t0 = tab, *t1 = tab + N[0], *t2 = tab + N[0] + N[1];
  long k = 0;
  for (long i0 = 0; i0 < N[0]; i0++)
    for (long i1 = 0; i1 < N[1]; i1++) {
      double p01 = t0[i0] * t1[i1];
      for (long i2 = 0; i2 < N[2]; i2++, k++) {
        double p = p01 * t2[i2];
        if (accumulate) { out[2 * k] += in[2 * k] * p; out[2 * k + 1] += in[2 * k + 1] * p; }
        else { out[2 * k] = in[2 * k] * p; out[2 * k + 1] = in[2 * k + 1] * p; }
      }
    }
}

/* ---------------------------------------------------------------------------
 * Matrix B on the ghost-padded grid (kernel/ndft-parallel.c:2703-3009 with the
 * stencil kernels of kernel/assign.c:478-1066).
 *   ngc[t] = no[t] + 2m; the node's lowest stencil index in the padded grid is
 *   u_t = floor(n_t x_t) - m - (-no_t/2) + m  (lowest_summation_index, :1563-1572).
 * Complex (c2c) and real (r2r, c2r plans) variants.
 * ------------------------------------------------------------------------- */
static void node_windows(int window, const long *n, const double *b, int m, const double *xj,
                         double *psi, double *dpsi, long *u, const long *no)
{
  const int c = 2 * m + 1;
  for (int t = 0; t < 3; t++) {
    double nx = (double)n[t] * xj[t];
    double fl = floor(nx);
    u[t] = (long)fl - m + no[t] / 2 + m;
    if (g_ip.order >= 0) {
      orc_intpol_axis(t, 0, m, nx, psi + c * t);
      if (dpsi) orc_intpol_axis(t, 1, m, nx, dpsi + c * t);
    } else {
      orc_psi_axis(window, n[t], b[t], m, nx, psi + c * t, dpsi ? dpsi + c * t : NULL);
    }
  }
}

void orc_B_trafo(const long *n, const long *no, int m, int window, const double *b, long M,
                 const double *x, const double *grid, int is_real, double *f, double *grad,
                 unsigned cflags)
{
  const int c = 2 * m + 1;
  const long g1 = no[1] + 2 * m, g2 = no[2] + 2 * m;
  const int want_f = (cflags & C_COMPUTE_F) && f;
  const int want_g = (cflags & C_COMPUTE_GRAD_F) && grad;
  const int cplx = is_real ? 1 : 2;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < M; j++) {
    double psi[3 * 64], dpsi[3 * 64];
    long u[3];
    node_windows(window, n, b, m, x + 3 * j, psi, want_g ? dpsi : NULL, u, no);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}; /* f, g0, g1, g2 (re, im) */
    for (int l0 = 0; l0 < c; l0++)
      for (int l1 = 0; l1 < c; l1++) {
        double pxy = psi[l0] * psi[c + l1];
        double pdxy = want_g ? dpsi[l0] * psi[c + l1] : 0.0;
        double pxdy = want_g ? psi[l0] * dpsi[c + l1] : 0.0;
        long row = ((u[0] + l0) * g1 + (u[1] + l1)) * g2 + u[2];
        for (int l2 = 0; l2 < c; l2++) {
          const double *gp = grid + cplx * (row + l2);
          double gr = gp[0], gi = is_real ? 0.0 : gp[1];
          double pz = psi[2 * c + l2];
          double wf = pxy * pz;
          acc[0] += wf * gr; acc[1] += wf * gi;
          if (want_g) {
            double w0 = pdxy * pz, w1 = pxdy * pz, w2 = pxy * dpsi[2 * c + l2];
            acc[2] += w0 * gr; acc[3] += w0 * gi;
            acc[4] += w1 * gr; acc[5] += w1 * gi;
            acc[6] += w2 * gr; acc[7] += w2 * gi;
          }
        }
      }
    /* assign_* accumulate into the outputs (*fv += f): assign.c:690, :792 */
    if (is_real) {
      if (want_f) f[j] += acc[0];
      if (want_g) for (int t = 0; t < 3; t++) grad[3 * j + t] += acc[2 + 2 * t];
    } else {
      if (want_f) { f[2 * j] += acc[0]; f[2 * j + 1] += acc[1]; }
      if (want_g) for (int t = 0; t < 3; t++) {
        grad[6 * j + 2 * t] += acc[2 + 2 * t]; grad[6 * j + 2 * t + 1] += acc[3 + 2 * t];
      }
    }
  }
}

/* Hessian gather (assign_hessian_f_c2c_pre_psi, kernel/assign.c:881-921, and
 * its r2r twin): hess[6j + c], c = xx, xy, xz, yy, yz, zz, added onto the
 * output as the reference's kernels do. */
void orc_B_trafo_hess(const long *n, const long *no, int m, int window, const double *b, long M,
                      const double *x, const double *grid, int is_real, double *hess)
{
  const int c = 2 * m + 1;
  const long g1 = no[1] + 2 * m, g2 = no[2] + 2 * m;
  const int cplx = is_real ? 1 : 2;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < M; j++) {
    double psi[3 * 64], dpsi[3 * 64], dd[3 * 64];
    long u[3];
    node_windows(window, n, b, m, x + 3 * j, psi, dpsi, u, no);
    for (int t = 0; t < 3; t++) {
      if (g_ip.order >= 0)
        orc_intpol_axis(t, 2, m, (double)n[t] * x[3 * j + t], dd + c * t);
      else
        ddpsi_axis(window, n[t], b[t], m, (double)n[t] * x[3 * j + t], psi + c * t, dpsi + c * t, dd + c * t);
    }
    double acc[12] = {0};
    for (int l0 = 0; l0 < c; l0++)
      for (int l1 = 0; l1 < c; l1++) {
        double xy = psi[l0] * psi[c + l1], dxy = dpsi[l0] * psi[c + l1], xdy = psi[l0] * dpsi[c + l1];
        double dxdy = dpsi[l0] * dpsi[c + l1], ddxy = dd[l0] * psi[c + l1], xddy = psi[l0] * dd[c + l1];
        long row = ((u[0] + l0) * g1 + (u[1] + l1)) * g2 + u[2];
        for (int l2 = 0; l2 < c; l2++) {
          const double *gp = grid + cplx * (row + l2);
          double gr = gp[0], gi = is_real ? 0.0 : gp[1];
          double z = psi[2 * c + l2], dz = dpsi[2 * c + l2], ddz = dd[2 * c + l2];
          double w[6] = {ddxy * z, dxdy * z, dxy * dz, xddy * z, xdy * dz, xy * ddz};
          for (int q = 0; q < 6; q++) { acc[2 * q] += w[q] * gr; acc[2 * q + 1] += w[q] * gi; }
        }
      }
    for (int q = 0; q < 6; q++) {
      if (is_real) hess[6 * j + q] += acc[2 * q];
      else { hess[12 * j + 2 * q] += acc[2 * q]; hess[12 * j + 2 * q + 1] += acc[2 * q + 1]; }
    }
  }
}

/* Scatter of one node into the ghost-padded grid (the inner loops of
 * kernel/assign.c:478-1066 for the adjoint). */
static void scatter_one(const long *n, const long *no, int m, int window, const double *b, long j,
                        const double *x, const double *f, const double *grad, int is_real, double *grid,
                        int use_f, int use_g)
{
  const int c = 2 * m + 1;
  const long g1 = no[1] + 2 * m, g2 = no[2] + 2 * m;
  const int cplx = is_real ? 1 : 2;
  double psi[3 * 64], dpsi[3 * 64];
  long u[3];
  node_windows(window, n, b, m, x + 3 * j, psi, use_g ? dpsi : NULL, u, no);
  double fr = 0, fi = 0, gv[6] = {0, 0, 0, 0, 0, 0};
  if (use_f) { fr = is_real ? f[j] : f[2 * j]; fi = is_real ? 0.0 : f[2 * j + 1]; }
  if (use_g) for (int t = 0; t < 3; t++) {
    gv[2 * t] = is_real ? grad[3 * j + t] : grad[6 * j + 2 * t];
    gv[2 * t + 1] = is_real ? 0.0 : grad[6 * j + 2 * t + 1];
  }
  for (int l0 = 0; l0 < c; l0++)
    for (int l1 = 0; l1 < c; l1++) {
      double pxy = psi[l0] * psi[c + l1];
      double pdxy = use_g ? dpsi[l0] * psi[c + l1] : 0.0;
      double pxdy = use_g ? psi[l0] * dpsi[c + l1] : 0.0;
      long row = ((u[0] + l0) * g1 + (u[1] + l1)) * g2 + u[2];
      for (int l2 = 0; l2 < c; l2++) {
        double *gp = grid + cplx * (row + l2);
        double pz = psi[2 * c + l2];
        double ar = 0, ai = 0;
        if (use_f) { double w = pxy * pz; ar += w * fr; ai += w * fi; }
        if (use_g) {
          double w0 = pdxy * pz, w1 = pxdy * pz, w2 = pxy * dpsi[2 * c + l2];
          ar += w0 * gv[0] + w1 * gv[2] + w2 * gv[4];
          ai += w0 * gv[1] + w1 * gv[3] + w2 * gv[5];
        }
        gp[0] += ar;
        if (!is_real) gp[1] += ai;
      }
    }
}

/* Adjoint B^T: scatter into the zeroed ghost-padded grid, node order as given
 * (sequential, so the accumulation order is the reference's unsorted order). */
void orc_B_adj(const long *n, const long *no, int m, int window, const double *b, long M,
               const double *x, const double *f, const double *grad, int is_real, double *grid,
               unsigned cflags)
{
  const int use_f = (cflags & C_COMPUTE_F) && f;
  const int use_g = (cflags & C_COMPUTE_GRAD_F) && grad;
  for (long j = 0; j < M; j++) scatter_one(n, no, m, window, b, j, x, f, grad, is_real, grid, use_f, use_g);
}

/* The same scatter on all OpenMP threads (CPU-baseline timing only).  Nodes
 * are grouped by the block (u0 / c, u1 / c) of their first stencil cell,
 * c = 2m+1; a node of block (s0, s1) writes padded planes [s0 c, s0 c + 2c - 1)
 * x [s1 c, s1 c + 2c - 1) only, so blocks whose indices have the same parity
 * in both axes never touch the same cell and run concurrently, four phases in
 * all.  Within a block the nodes keep their input order. */
void orc_B_adj_blocks(const long *n, const long *no, int m, int window, const double *b, long M,
                      const double *x, const double *f, const double *grad, int is_real, double *grid,
                      unsigned cflags)
{
  const int c = 2 * m + 1;
  const int use_f = (cflags & C_COMPUTE_F) && f;
  const int use_g = (cflags & C_COMPUTE_GRAD_F) && grad;
  const long nb0 = (no[0] + 2 * m + c - 1) / c, nb1 = (no[1] + 2 * m + c - 1) / c;
  const long nblk = nb0 * nb1;
  long *blk = malloc(sizeof(long) * (M > 0 ? M : 1));
  long *start = calloc((size_t)nblk + 1, sizeof(long));
  long *order = malloc(sizeof(long) * (M > 0 ? M : 1));
  for (long j = 0; j < M; j++) {
    long u[2];
    for (int t = 0; t < 2; t++) u[t] = (long)floor((double)n[t] * x[3 * j + t]) - m + no[t] / 2 + m;
    blk[j] = (u[0] / c) * nb1 + u[1] / c;
    start[blk[j] + 1]++;
  }
  for (long k = 0; k < nblk; k++) start[k + 1] += start[k];
  {
    long *fill = malloc(sizeof(long) * (size_t)(nblk > 0 ? nblk : 1));
    memcpy(fill, start, sizeof(long) * (size_t)nblk);
    for (long j = 0; j < M; j++) order[fill[blk[j]]++] = j;
    free(fill);
  }
  for (int ph = 0; ph < 4; ph++) {
    const long p0 = ph >> 1, p1 = ph & 1;
    const long h0 = (nb0 - p0 + 1) / 2, h1 = (nb1 - p1 + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (long q = 0; q < h0 * h1; q++) {
      const long k = (p0 + 2 * (q / h1)) * nb1 + p1 + 2 * (q % h1);
      for (long i = start[k]; i < start[k + 1]; i++)
        scatter_one(n, no, m, window, b, order[i], x, f, grad, is_real, grid, use_f, use_g);
    }
  }
  free(blk);
  free(start);
  free(order);
}

/* ---------------------------------------------------------------------------
 * Node sort (kernel/ndft-parallel.c:2121-2159 with util/util.c:200-280):
 * key_j = row-major index of ((floor(n_t x_t - m) mod n_t)_t), stable LSD
 * radix sort with 9-bit digits over rhigh = ceil(log2(prod n)) - 1 bits.
 * out[2p] = key, out[2p+1] = node index (the reference's ar_x layout).
 * ------------------------------------------------------------------------- */
void orc_sort_nodes(const long *n, int m, long M, const double *x, long *out)
{
  long *tmp = malloc(sizeof(long) * 2 * (M > 0 ? M : 1));
  for (long i = 0; i < M; i++) {
    long key = 0;
    for (int t = 0; t < 3; t++) {
      long h = (long)floor((double)n[t] * x[3 * i + t] - m);
      long u = (h % n[t] + n[t]) % n[t];
      key += u;
      if (t + 1 < 3) key *= n[t + 1];
    }
    out[2 * i] = key;
    out[2 * i + 1] = i;
  }
  double nprod = (double)n[0] * (double)n[1] * (double)n[2];
  long rhigh = (long)ceil(log2(nprod)) - 1;
  const long rwidth = 9, radix_n = 1 << 9, mask = radix_n - 1;
  long *from = out, *to = tmp;
  long shift = 0;
  long counts[512];
  while (rhigh >= 0) {
    memset(counts, 0, sizeof(counts));
    for (long i = 0; i < M; i++) counts[(from[2 * i] >> shift) & mask]++;
    long acc = 0;
    for (long d = 0; d < radix_n; d++) { long c = counts[d]; counts[d] = acc; acc += c; }
    for (long i = 0; i < M; i++) {
      long d = (from[2 * i] >> shift) & mask;
      long p = counts[d]++;
      to[2 * p] = from[2 * i];
      to[2 * p + 1] = from[2 * i + 1];
    }
    long *sw = from; from = to; to = sw;
    shift += rwidth;
    rhigh -= rwidth;
  }
  if (from != out) memcpy(out, from, sizeof(long) * 2 * M);
  free(tmp);
}
