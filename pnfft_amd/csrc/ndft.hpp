// ndft.hpp -- host interface of the direct NDFT kernels (ndft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace pnfft_amd {

template <class R>
struct NdftArgs {
  int N[3];
  long M;
  const R* x;
  Cplx<R>* fhat;
  Cplx<R>* f;
  Cplx<R>* grad;
  Cplx<R>* hess;  // C[6M] (R[6M] for c2r), components xx, xy, xz, yy, yz, zz
  int want_f;
  int want_grad;
  int want_hess;
  int c2r;  // real plan: f_hat is the half spectrum k2 in [-N2/2, 0]; f, grad are real arrays
};

template <class R>
int ndft_trafo(const NdftArgs<R>& A, hipStream_t s);

template <class R>
int ndft_adj(const NdftArgs<R>& A, hipStream_t s);

}  // namespace pnfft_amd
