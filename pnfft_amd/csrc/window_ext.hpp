// window_ext.hpp -- windows that the B kernels take only from per-node tables:
// the Bessel-I0 window (kernel/ndft-parallel.c:2195-2237) and the interpolated
// windows of PNFFT_PRE_{CONST,LIN,QUAD,CUB}_PSI (kernel/ndft-parallel.c:1586-1617,
// kernel/ipnfft.h:457-497).  They are evaluated by the table kernels
// (window_tab.hip, and the Hessian's table kernel), never inside the gather and
// scatter kernels, whose window evaluation stays that of the other windows.
#pragma once

#include "window_math.hpp"

namespace pnfft_amd {

// Interpolation tables of a plan: for axis t and derivative d, K (2m+1)
// (order+1) values psi^(d)((m + (k+i)/K - c)/n_t) in the order (k, c, i),
// i = -order/2 .. (order+1)/2 (init_intpol_table_psi, kernel/ndft-parallel.c:318-352).
template <class R>
struct IntpolTabs {
  int order = -1;  // -1: no interpolation
  long K = 0;      // interpolation nodes per unit interval
  const R* tab[3][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
};

// I0 and I1 by their power series, summed for a fixed 96 terms: the terms
// q^k / (k!)^2 (q = x^2/4) fall below 1e-30 of the sum by then for x <= 52,
// which covers b m of every plan (b <= 2 pi, m <= 8).
__host__ __device__ inline double series_i0(double x) {
  const double q = 0.25 * x * x;
  double t = 1.0, s = 1.0;
  for (int k = 1; k <= 96; k++) {
    t *= q / (double(k) * double(k));
    s += t;
  }
  return s;
}

__host__ __device__ inline double series_i1(double x) {
  const double q = 0.25 * x * x;
  double t = 0.5 * x, s = t;
  for (int k = 1; k <= 96; k++) {
    t *= q / (double(k) * double(k + 1));
    s += t;
  }
  return s;
}

// interpolation of order 0..3 at dist in [0, 1) from f[0..order] (kernel/ipnfft.h:457-497)
template <class R>
__host__ __device__ inline R intpol_eval(int order, const R* f, R d) {
  switch (order) {
    case 0: return f[0];
    case 1: return f[0] * (R(1) - d) + f[1] * d;
    case 2: {
      const R c0 = d + R(1), c1 = d, c2 = d - R(1);
      return R(0.5) * f[0] * c1 * c2 - c0 * f[1] * c2 + R(0.5) * c0 * c1 * f[2];
    }
    default: {
      const R c0 = d + R(1), c1 = d, c2 = d - R(1), c3 = d - R(2);
      return (-f[0] * c1 * c2 * c3 + R(3) * c0 * f[1] * c2 * c3 - R(3) * c0 * c1 * f[2] * c3 + c0 * c1 * c2 * f[3]) /
             R(6);
    }
  }
}

// psi (and psi', psi'' if WD / WDD) on the stencil of axis t for fractional
// part frac: interpolated (I.order >= 0), the Bessel-I0 window evaluated in
// FP64, or the standard windows of window_math.hpp.
template <class R, int CM, bool WD, bool WDD>
__device__ inline void ext_window_axis(const GridGeom& g, const IntpolTabs<R>& I, int t, R frac, R (&psi)[CM],
                                       R (&dpsi)[CM], R (&dd)[CM]) {
  constexpr int m = (CM - 1) / 2;
  if (I.order >= 0) {
    // pre_tensor_intpol: dist K in the compute precision, k clamped into the table
    const R dk_full = frac * (R)I.K;
    long k = (long)floor(dk_full);
    if (k < 0) k = 0;
    if (k > I.K - 1) k = I.K - 1;
    const R dk = dk_full - (R)k;
    const int w = I.order + 1;
#pragma unroll
    for (int s = 0; s < CM; s++) {
      const long base = ((long)k * CM + s) * w;
      psi[s] = intpol_eval<R>(I.order, I.tab[0][t] + base, dk);
      if (WD) dpsi[s] = intpol_eval<R>(I.order, I.tab[1][t] + base, dk);
      if (WDD) dd[s] = intpol_eval<R>(I.order, I.tab[2][t] + base, dk);
    }
    return;
  }
  if (g.win == WIN_BESSEL_I0) {
    const double n = g.n[t], b = g.b[t];
#pragma unroll
    for (int s = 0; s < CM; s++) {
      const double xn = (double)frac + double(m - s);
      const double d = double(m) * m - xn * xn;
      double v = 0, v1 = 0, v2 = 0;
      if (d > 0) {
        const double r = sqrt(d), i0 = series_i0(b * r), i1 = series_i1(b * r);
        v = 0.5 * i0;
        v1 = -0.5 * b * n * xn * i1 / r;
        v2 = 0.5 * b * n * n / d * (b * xn * xn * i0 - i1 / r * (xn * xn + double(m) * m));
      } else if (d == 0) {
        v = 0.5;
        v1 = -(0.5 * b * n) * (0.5 * b * n) * xn / n;
        const double q = b * b * m * n, h = b * n;
        v2 = q * q / 16.0 - h * h / 4.0;
      }
      psi[s] = (R)v;
      if (WD) dpsi[s] = (R)v1;
      if (WDD) dd[s] = (R)v2;
    }
    return;
  }
  window_axis_s<R, CM, WD || WDD>(g.win, frac, (R)g.n[t], (R)g.b[t], psi, dpsi);
  if (WDD) window_axis_dd_s<R, CM>(g.win, frac, (R)g.n[t], (R)g.b[t], psi, dpsi, dd);
}

}  // namespace pnfft_amd
