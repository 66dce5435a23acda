// b_common.hpp -- per-node stencil geometry shared by the B kernels and the binning.
#pragma once

#include "device_common.hpp"

namespace pnfft_amd {

// Natural grid index of stencil point s = 0 on one axis, (floor(n x) - m) mod n
// (kernel/ndft-parallel.c:2165-2174), and the FP64 fractional part of n x.
template <class R>
struct NodeAxis {
  int start;  // natural index of stencil point s = 0
  R frac;     // n x - floor(n x)
};

// n is a power of two (plans reject other oversampled lengths): floor(n x) mod n
// is fl - n floor(fl / n), exact in FP64 for |n x| < 2^53 (the scaling by 1/n is
// an exponent shift), and the shift by -m wraps with a mask.  sh (0 or 1/2, the
// interlaced pass) is added to n x in FP64, where both are exact.
template <class R>
__device__ inline NodeAxis<R> node_axis(R xv, int n, int m, double sh) {
  const double nx = (double)n * (double)xv + sh;
  const double fl = floor(nx);
  NodeAxis<R> a;
  a.frac = (R)(nx - fl);
  const double q = floor(ldexp(fl, -__builtin_ctz((unsigned)n)));
  const int r = (int)fma(-q, (double)n, fl);
  a.start = (r - m) & (n - 1);
  return a;
}

}  // namespace pnfft_amd
