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

template <class R>
__device__ inline NodeAxis<R> node_axis(R xv, int n, int m) {
  const double nx = (double)n * (double)xv;
  const double fl = floor(nx);
  NodeAxis<R> a;
  a.frac = (R)(nx - fl);
  long st = ((long)fl - m) % n;
  if (st < 0) st += n;
  a.start = (int)st;
  return a;
}

}  // namespace pnfft_amd
