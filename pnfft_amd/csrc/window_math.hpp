// window_math.hpp -- window functions psi / psi' on the (2m+1)-point stencil
// of one axis, for host and device, templated on the compute precision.
//
// For a node coordinate x on an axis with oversampled length n, let
// fl = floor(n x) (taken in FP64) and frac = n x - fl in [0, 1).  Stencil
// point s = 0..2m is grid index l_s = fl - m + s and
//   psi[s]  = psi(x - l_s/n),   dpsi[s] = psi'(x - l_s/n)  (d/dx).
// Window formulas follow kernel/ndft-parallel.c:1678-1797 (psi) and
// :1850-1953 (psi'), with the 1-d kernels of :2241-2269 and the sinc of
// kernel/sinc.c:26-52.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace pnfft_amd {

enum WindowId : int { WIN_KB = 0, WIN_GAUSS = 1, WIN_BSPLINE = 2, WIN_SINCPOW = 3, WIN_BESSEL_I0 = 4 };

constexpr int kMaxM = 16;                 // largest cutoff the window formulas handle (host)
constexpr int kMaxC = 2 * kMaxM + 1;      // largest stencil width (host)
constexpr int kMinPlanM = 2;              // cutoffs the B kernels are compiled for
constexpr int kMaxPlanM = 8;

template <class R> struct Consts;
template <> struct Consts<double> {
  // PNFFT_EPSILON of the double library (kernel/ipnfft.h:99)
  static constexpr double eps = 2.220446049250313080847e-16;
};
template <> struct Consts<float> {
  // PNFFT_EPSILON of the single library is LDBL_EPSILON (kernel/ipnfft.h:81)
  static constexpr float eps = 1.08420217248550443e-19f;
};

__host__ __device__ inline double pi_v(double) { return 3.14159265358979323846; }
__host__ __device__ inline float pi_v(float) { return 3.14159265358979323846f; }

__host__ __device__ inline double m_sqrt(double v) { return sqrt(v); }
__host__ __device__ inline float m_sqrt(float v) { return sqrtf(v); }
__host__ __device__ inline double m_sin(double v) { return sin(v); }
__host__ __device__ inline float m_sin(float v) { return sinf(v); }
__host__ __device__ inline double m_cos(double v) { return cos(v); }
__host__ __device__ inline float m_cos(float v) { return cosf(v); }
__host__ __device__ inline double m_sinh(double v) { return sinh(v); }
__host__ __device__ inline float m_sinh(float v) { return sinhf(v); }
__host__ __device__ inline double m_cosh(double v) { return cosh(v); }
__host__ __device__ inline float m_cosh(float v) { return coshf(v); }
__host__ __device__ inline double m_exp(double v) { return exp(v); }
__host__ __device__ inline float m_exp(float v) { return expf(v); }
__host__ __device__ inline double m_tan(double v) { return tan(v); }
__host__ __device__ inline float m_tan(float v) { return tanf(v); }
__host__ __device__ inline double m_fabs(double v) { return fabs(v); }
__host__ __device__ inline float m_fabs(float v) { return fabsf(v); }

// sin(x)/x with the small-argument branches of kernel/sinc.c:26-52.
template <class R>
__host__ __device__ inline R sinc_ref(R x) {
  const R b = Consts<R>::eps;
  const R bs = m_sqrt(b);
  const R bs2 = m_sqrt(bs);
  if (m_fabs(x) >= bs2) return m_sin(x) / x;
  R r = R(1);
  if (m_fabs(x) >= b) {
    const R x2 = x * x;
    r -= x2 / R(6);
    if (m_fabs(x) >= bs) r += (x2 * x2) / R(120);
  }
  return r;
}

// Integer power by repeated squaring (the exponent 2m of the sinc-power window).
template <class R>
__host__ __device__ inline R ipow(R v, int e) {
  R r = R(1);
  while (e > 0) {
    if (e & 1) r *= v;
    v *= v;
    e >>= 1;
  }
  return r;
}

// V[d] = M_k(u + d), d = 0..k-1, for the cardinal B-spline M_k on [0, k] and
// u in (0, 1]; recursion M_r(t) = (t M_{r-1}(t) + (r - t) M_{r-1}(t-1))/(r-1).
template <class R>
__host__ __device__ inline void bspline_values(int k, R u, R* V) {
  V[0] = R(1);
  for (int r = 2; r <= k; r++) {
    const R inv = R(1) / R(r - 1);
    V[r - 1] = R(0);
    for (int d = r - 1; d >= 0; d--) {
      const R left = (d > 0) ? V[d - 1] : R(0);
      V[d] = ((u + R(d)) * V[d] + (R(r) - u - R(d)) * left) * inv;
    }
  }
}

// (sinh z - z cosh z) / z^3 and (sin z - z cos z) / z^3 without cancellation.
template <class R>
__host__ __device__ inline R kb_h_sinh(R z) {
  const R z2 = z * z;
  if (z2 < R(0.25)) {
    // -sum_{k>=1} 2k z^{2k-2} / (2k+1)!
    R term = R(1), sum = R(0);
    R fact = R(6);  // (2k+1)! for k = 1
    for (int k = 1; k <= 8; k++) {
      sum += R(2 * k) * term / fact;
      term *= z2;
      fact *= R((2 * k + 2) * (2 * k + 3));
    }
    return -sum;
  }
  return (m_sinh(z) - z * m_cosh(z)) / (z2 * z);
}

template <class R>
__host__ __device__ inline R kb_h_sin(R z) {
  const R z2 = z * z;
  if (z2 < R(0.25)) {
    // sum_{k>=1} (-1)^{k+1} 2k z^{2k-2} / (2k+1)!
    R term = R(1), sum = R(0), sgn = R(1);
    R fact = R(6);
    for (int k = 1; k <= 8; k++) {
      sum += sgn * R(2 * k) * term / fact;
      term *= z2;
      sgn = -sgn;
      fact *= R((2 * k + 2) * (2 * k + 3));
    }
    return sum;
  }
  return (m_sin(z) - z * m_cos(z)) / (z2 * z);
}

// Kaiser-Bessel psi at grid offset u = l - n x (kaiser_bessel_1d, :2241-2254).
template <class R>
__host__ __device__ inline R kb_psi(R u, R b, int m) {
  const R d = R(m) * R(m) - u * u;
  const R r = m_sqrt(m_fabs(d));
  const R pi = pi_v(R(0));
  if (d < R(0)) return m_sin(b * r) / (pi * r);
  return (d > R(0)) ? m_sinh(b * r) / (pi * r) : b / pi;
}

// kb_psi for |u| <= m (d >= 0): the sinh branch and the edge value only.  For
// u = (s - m) - frac with s >= 1 and frac in [0, 1], |u| <= m holds in floating
// point as well (the exact value does not exceed m, a representable number),
// so a stencil's points s >= 1 never take the sin branch.
template <class R>
__host__ __device__ inline R kb_psi_in(R u, R b, int m) {
  const R d = R(m) * R(m) - u * u;
  const R r = m_sqrt(d);
  const R pi = pi_v(R(0));
  return (d > R(0)) ? m_sinh(b * r) / (pi * r) : b / pi;
}

#if defined(__HIP_DEVICE_COMPILE__)
// FP32 Kaiser-Bessel at a stencil point with |u| <= m (the points s >= 1, see
// kb_psi_in), psi and psi' from one exponential.  With r = sqrt(m^2 - u^2) and
// z = b r:  psi = (b / pi) S(z), S(z) = sinh z / z, and psi' = n xn b^3 H(z) / pi,
// H(z) = (sinh z - z cosh z) / z^3, xn = -u.  For z >= 1, E = e^z gives
// S = (E - 1/E) / (2 z) and H = ((E - 1/E) - z (E + 1/E)) / (2 z^3); below 1
// their Taylor series S = sum_k z^2k / (2k+1)!, H = -sum_{k>=1} 2k z^(2k-2) / (2k+1)!
// (truncated after z^10, remainder < 2e-10).  The d = 0 edge takes the reference's
// value -n m b^3 / (3 pi) for psi'.
__device__ inline void kb_in_f32(float u, float b, int m, float nR, float& psi, float* dpsi) {
  const float d = float(m) * float(m) - u * u;
  const float r = __builtin_amdgcn_sqrtf(d);
  const float z = b * r, z2 = z * z;
  const float E = expf(z), Ei = __builtin_amdgcn_rcpf(E), iz = __builtin_amdgcn_rcpf(z);
  const bool small = z < 1.f;
  const float Ss = 1.f + z2 * (1.f / 6.f + z2 * (1.f / 120.f + z2 * (1.f / 5040.f + z2 * (1.f / 362880.f +
                                                                                            z2 * (1.f / 39916800.f)))));
  const float S = small ? Ss : 0.5f * (E - Ei) * iz;
  const float bpi = b * 0.318309886183790671538f;  // b / pi
  psi = bpi * S;
  if (dpsi) {
    const float Hs = -(2.f / 6.f + z2 * (4.f / 120.f + z2 * (6.f / 5040.f + z2 * (8.f / 362880.f +
                                                                                    z2 * (10.f / 39916800.f + z2 * (12.f / 6227020800.f))))));
    const float H = small ? Hs : 0.5f * ((E - Ei) - z * (E + Ei)) * (iz * iz * iz);
    *dpsi = (d == 0.f) ? -nR * float(m) * b * b * b / (3.f * 3.14159265358979323846f) : nR * (-u) * b * b * bpi * H;
  }
}
#endif

// Kaiser-Bessel psi' at xn = n x - l (kaiser_bessel_derivative_1d, :2256-2269).
// FP64 keeps the reference's formula; FP32 uses the cancellation-free form
// n xn b^3 h(b r)/pi.  The d == 0 value is the reference's -n m b^3/(3 pi) on
// both stencil edges.
template <class R>
__host__ __device__ inline R kb_dpsi(R xn, R nR, R b, int m, R psi) {
  const R d = R(m) * R(m) - xn * xn;
  const R r = m_sqrt(m_fabs(d));
  const R pi = pi_v(R(0));
  if (d == R(0)) return -nR * R(m) * b * b * b / (R(3) * pi);
  if (sizeof(R) == sizeof(double)) {
    if (d < R(0)) return nR * xn / d * (psi - b * m_cos(b * r) / pi);
    return nR * xn / d * (psi - b * m_cosh(b * r) / pi);
  }
  const R z = b * r;
  if (d > R(0)) return nR * xn * b * b * b * kb_h_sinh(z) / pi;
  return -nR * xn * b * b * b * kb_h_sin(z) / pi;
}

// kb_dpsi for |xn| <= m (d >= 0), as kb_psi_in: xn = frac + (m - s) with s >= 1
template <class R>
__host__ __device__ inline R kb_dpsi_in(R xn, R nR, R b, int m, R psi) {
  const R d = R(m) * R(m) - xn * xn;
  const R r = m_sqrt(d);
  const R pi = pi_v(R(0));
  if (d == R(0)) return -nR * R(m) * b * b * b / (R(3) * pi);
  if (sizeof(R) == sizeof(double)) return nR * xn / d * (psi - b * m_cosh(b * r) / pi);
  return nR * xn * b * b * b * kb_h_sinh(b * r) / pi;
}

// Window values of one axis.  frac in [0,1), nR = n as R.  dpsi may be null.
template <class R>
__host__ __device__ inline void window_axis(int win, int m, R frac, R nR, R b, R* psi, R* dpsi) {
  const int c = 2 * m + 1;
  const R pi = pi_v(R(0));
  switch (win) {
    case WIN_GAUSS: {
      const R norm = R(1) / m_sqrt(pi * b);
      for (int s = 0; s < c; s++) {
        const R u = R(s - m) - frac;
        psi[s] = m_exp(-(u * u) / b) * norm;
      }
      if (dpsi)
        for (int s = 0; s < c; s++) dpsi[s] = R(-2) * nR / b * (frac + R(m - s)) * psi[s];
      break;
    }
    case WIN_BSPLINE: {
      R V[2 * kMaxM], W[2 * kMaxM];
      const R u = R(1) - frac;
      bspline_values<R>(2 * m, u, V);
      psi[0] = R(0);
      for (int s = 1; s < c; s++) psi[s] = V[s - 1];
      if (dpsi) {
        bspline_values<R>(2 * m - 1, u, W);
        for (int s = 0; s < c; s++) {
          const R a = (s >= 1 && s - 1 < 2 * m - 1) ? W[s - 1] : R(0);
          const R bb = (s >= 2 && s - 2 < 2 * m - 1) ? W[s - 2] : R(0);
          dpsi[s] = -nR * (a - bb);
        }
      }
      break;
    }
    case WIN_SINCPOW: {
      for (int s = 0; s < c; s++) {
        const R u = R(s - m) - frac;
        psi[s] = ipow(sinc_ref<R>(pi * u / b), 2 * m) / b;
      }
      if (dpsi)
        for (int s = 0; s < c; s++) {
          const R y = pi * (frac + R(m - s)) / b;
          dpsi[s] = (m_fabs(y) > Consts<R>::eps)
                        ? R(2 * m) * pi * nR / b * (R(1) / m_tan(y) - R(1) / y) * psi[s]
                        : R(0);
        }
      break;
    }
    default: {  // Kaiser-Bessel
      for (int s = 0; s < c; s++) psi[s] = kb_psi<R>(R(s - m) - frac, b, m);
      if (dpsi)
        for (int s = 0; s < c; s++) dpsi[s] = kb_dpsi<R>(frac + R(m - s), nR, b, m, psi[s]);
      break;
    }
  }
}

// cot y - 1/y and 1/y^2 - 1 - cot^2 y, the factors of the sinc-power window's
// psi' and psi'' (kernel/ndft-parallel.c:1907-1915, :2052-2071).  FP32 takes
// them from their Taylor series for |y| < 1/2, where the closed forms cancel
// (-y/3 - y^3/45 - 2y^5/945 - ..., -1/3 - y^2/15 - 2y^4/189 - ...); FP64 keeps
// the reference's closed forms.
template <class R>
__host__ __device__ inline R sincpow_c1(R y) {
  if (sizeof(R) == sizeof(float) && m_fabs(y) < R(0.5)) {
    const R y2 = y * y;
    return -y * (R(1.0 / 3) +
                 y2 * (R(1.0 / 45) +
                       y2 * (R(2.0 / 945) + y2 * (R(1.0 / 4725) + y2 * (R(2.0 / 93555) + y2 * R(1382.0 / 638512875))))));
  }
  return R(1) / m_tan(y) - R(1) / y;
}

template <class R>
__host__ __device__ inline R sincpow_c2(R y) {
  if (sizeof(R) == sizeof(float) && m_fabs(y) < R(0.5)) {
    const R y2 = y * y;
    return -(R(1.0 / 3) +
             y2 * (R(1.0 / 15) +
                   y2 * (R(2.0 / 189) +
                         y2 * (R(1.0 / 675) +
                               y2 * (R(2.0 / 10395) + y2 * (R(1382.0 / 58046625) + y2 * R(4.0 / 1403325)))))));
  }
  const R t = m_tan(y);
  return R(1) / (y * y) - R(1) - R(1) / (t * t);
}

// bspline_values with the order K fixed at compile time: same recursion and
// operation order, every index static so V stays in registers.
// one level r of the recursion below: the r values of order r from the r - 1
// values of order r - 1 in V[0, r - 1), in place
template <class R>
__host__ __device__ inline void bspline_level(R u, int r, R* V) {
  const R inv = R(1) / R(r - 1);
  V[r - 1] = R(0);
#pragma unroll
  for (int d = r - 1; d >= 0; d--) {
    const R left = (d > 0) ? V[d - 1] : R(0);
    V[d] = ((u + R(d)) * V[d] + (R(r) - u - R(d)) * left) * inv;
  }
}

template <int K, class R>
__host__ __device__ inline void bspline_values_s(R u, R (&V)[K]) {
  V[0] = R(1);
#pragma unroll
  for (int r = 2; r <= K; r++) bspline_level(u, r, V);
}

// window_axis for a stencil width CM = 2m+1 known at compile time (the device
// kernels for m = 2..8).  Values are identical to window_axis.
template <class R, int CM, bool WD>
__host__ __device__ inline void window_axis_s(int win, R frac, R nR, R b, R (&psi)[CM], R (&dpsi)[CM]) {
  constexpr int m = (CM - 1) / 2;
  const R pi = pi_v(R(0));
  switch (win) {
    case WIN_GAUSS: {
      const R norm = R(1) / m_sqrt(pi * b);
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const R u = R(s - m) - frac;
        psi[s] = m_exp(-(u * u) / b) * norm;
      }
      if (WD) {
#pragma unroll
        for (int s = 0; s < CM; s++) dpsi[s] = R(-2) * nR / b * (frac + R(m - s)) * psi[s];
      }
      break;
    }
    case WIN_BSPLINE: {
      R V[2 * m];
      const R u = R(1) - frac;
      R W[2 * m - 1];
      if (WD) {
        // the order 2m - 1 values of the derivative are the recursion's last
        // level but one: one more level gives the window values
        bspline_values_s<2 * m - 1, R>(u, W);
#pragma unroll
        for (int d = 0; d < 2 * m - 1; d++) V[d] = W[d];
        bspline_level(u, 2 * m, V);
      } else {
        bspline_values_s<2 * m, R>(u, V);
      }
      psi[0] = R(0);
#pragma unroll
      for (int s = 1; s < CM; s++) psi[s] = V[s - 1];
      if (WD) {
#pragma unroll
        for (int s = 0; s < CM; s++) {
          const R a = (s >= 1 && s - 1 < 2 * m - 1) ? W[s - 1] : R(0);
          const R bb = (s >= 2 && s - 2 < 2 * m - 1) ? W[s - 2] : R(0);
          dpsi[s] = -nR * (a - bb);
        }
      }
      break;
    }
    case WIN_SINCPOW: {
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const R u = R(s - m) - frac;
        psi[s] = ipow(sinc_ref<R>(pi * u / b), 2 * m) / b;
      }
      if (WD) {
#pragma unroll
        for (int s = 0; s < CM; s++) {
          const R y = pi * (frac + R(m - s)) / b;
          dpsi[s] = (m_fabs(y) > Consts<R>::eps) ? R(2 * m) * pi * nR / b * sincpow_c1(y) * psi[s] : R(0);
        }
      }
      break;
    }
    default: {  // Kaiser-Bessel: only the point s = 0 can lie outside |u| <= m
      psi[0] = kb_psi<R>(R(-m) - frac, b, m);
      if (WD) dpsi[0] = kb_dpsi<R>(frac + R(m), nR, b, m, psi[0]);
#if defined(__HIP_DEVICE_COMPILE__)
      if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int s = 1; s < CM; s++) kb_in_f32(R(s - m) - frac, b, m, nR, psi[s], WD ? &dpsi[s] : nullptr);
        break;
      }
#endif
#pragma unroll
      for (int s = 1; s < CM; s++) psi[s] = kb_psi_in<R>(R(s - m) - frac, b, m);
      if (WD) {
#pragma unroll
        for (int s = 1; s < CM; s++) dpsi[s] = kb_dpsi_in<R>(frac + R(m - s), nR, b, m, psi[s]);
      }
      break;
    }
  }
}

// k(z) = h'(z)/z for the two KB branches, h as in kb_h_sinh / kb_h_sin:
// (3 z cosh z - (z^2 + 3) sinh z) / z^5 and (z^2 sin z - 3 sin z + 3 z cos z) / z^5,
// from their series -sum_{k>=2} (+-1)^k 2k (2k-2) z^{2k-4} / (2k+1)! where the
// closed forms cancel (z^2 < 4).
template <class R>
__host__ __device__ inline R kb_k_sinh(R z) {
  const R z2 = z * z;
  if (z2 < R(4)) {
    R term = R(1), sum = R(0);
    R fact = R(120);  // (2k+1)! for k = 2
    for (int k = 2; k <= 12; k++) {
      sum += R(2 * k * (2 * k - 2)) * term / fact;
      term *= z2;
      fact *= R((2 * k + 2) * (2 * k + 3));
    }
    return -sum;
  }
  return (R(3) * z * m_cosh(z) - (z2 + R(3)) * m_sinh(z)) / (z2 * z2 * z);
}

template <class R>
__host__ __device__ inline R kb_k_sin(R z) {
  const R z2 = z * z;
  if (z2 < R(4)) {
    R term = R(1), sum = R(0), sgn = R(-1);
    R fact = R(120);
    for (int k = 2; k <= 12; k++) {
      sum += sgn * R(2 * k * (2 * k - 2)) * term / fact;
      term *= z2;
      sgn = -sgn;
      fact *= R((2 * k + 2) * (2 * k + 3));
    }
    return sum;
  }
  return (z2 * m_sin(z) - R(3) * m_sin(z) + R(3) * z * m_cos(z)) / (z2 * z2 * z);
}

// Kaiser-Bessel psi'' at xn = n x - l (kaiser_bessel_second_derivative_1d,
// kernel/ndft-parallel.c:2272-2285).  FP64 keeps the reference's formula; FP32
// uses n^2 b^3/pi (h(z) - b^2 xn^2 k(z)) (z = b r), its cancellation-free form,
// and -n^2 b^3/pi (h_s(z) + b^2 xn^2 k_s(z)) outside the support.
template <class R>
__host__ __device__ inline R kb_ddpsi(R xn, R nR, R b, int m, R psi, R dpsi) {
  const R d = R(m) * R(m) - xn * xn;
  const R r = m_sqrt(m_fabs(d));
  const R pi = pi_v(R(0));
  if (d == R(0)) {
    const R bn = b * nR, bm = b * R(m);
    return b * bn * bn / (R(15) * pi) * (bm * bm - R(5));
  }
  if (sizeof(R) == sizeof(double)) {
    const R hyp = (d < R(0)) ? m_cos(b * r) : m_cosh(b * r);
    return R(3) * nR * xn * dpsi / d + nR * nR * psi / d * (R(1) + b * xn * b * xn) - b * nR * nR / (pi * d) * hyp;
  }
  const R z = b * r, c = nR * nR * b * b * b / pi, bx2 = b * b * xn * xn;
  if (d > R(0)) return c * (kb_h_sinh(z) - bx2 * kb_k_sinh(z));
  return -c * (kb_h_sin(z) + bx2 * kb_k_sin(z));
}

// psi'' on the stencil of one axis (pre_ddpsi_tensor_*, kernel/ndft-parallel.c:
// 1956-2105), from psi and psi' of the same node; width CM = 2m+1 static.
template <class R, int CM>
__host__ __device__ inline void window_axis_dd_s(int win, R frac, R nR, R b, const R (&psi)[CM],
                                                 const R (&dpsi)[CM], R (&dd)[CM]) {
  constexpr int m = (CM - 1) / 2;
  const R pi = pi_v(R(0));
  const R nn = nR * nR;
  switch (win) {
    case WIN_GAUSS: {
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const R u = frac + R(m - s);
        dd[s] = R(2) * nn / b * (R(2) / b * u * u - R(1)) * psi[s];
      }
      break;
    }
    case WIN_BSPLINE: {
      // n^2 (M_{2m-2}(t) - 2 M_{2m-2}(t-1) + M_{2m-2}(t-2)), Q[d] = M_{2m-2}(1 - frac + d)
      R Q[2 * m - 2];
      bspline_values_s<2 * m - 2, R>(R(1) - frac, Q);
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const R q3 = (s >= 3 && s - 3 <= 2 * m - 3) ? Q[s - 3] : R(0);
        const R q2 = (s >= 2 && s - 2 <= 2 * m - 3) ? Q[s - 2] : R(0);
        const R q1 = (s >= 1 && s - 1 <= 2 * m - 3) ? Q[s - 1] : R(0);
        dd[s] = nn * (q3 - R(2) * q2 + q1);
      }
      break;
    }
    case WIN_SINCPOW: {
      const R a = pi * nR / b;
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const R y = pi * (frac + R(m - s)) / b;
        if (m_fabs(y) > Consts<R>::eps) {
          dd[s] = R(2 * m) * a * sincpow_c1(y) * dpsi[s] + R(2 * m) * a * a * sincpow_c2(y) * psi[s];
        } else {
          dd[s] = -R(2 * m) * a * a / (R(3) * b);
        }
      }
      break;
    }
    default: {  // Kaiser-Bessel
#pragma unroll
      for (int s = 0; s < CM; s++) dd[s] = kb_ddpsi<R>(frac + R(m - s), nR, b, m, psi[s], dpsi[s]);
      break;
    }
  }
}

}  // namespace pnfft_amd
