// fft.hpp -- host interface of the pruned 3-D FFT passes (fft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"

namespace pnfft_amd {

// Lengths handled by the LDS pencil kernel: powers of two up to 8192 (FP32)
// or 4096 (FP64).
bool fft_length_supported(int L, bool is_double);

template <class R>
struct FftAxis {
  int L = 0;
  int C = 0;          // pencils per block, rows pass
  int Ccols = 0;         // pencils per block, axis-0 columns pass
  int Ccols_narrow = 0;  // pencils per block, axis-1 columns pass
  Cplx<R>* tw = nullptr;   // device twiddles e^{-2 pi i j / L}
  Cplx<R>* tw2 = nullptr;  // c2r axis 2: e^{-i pi j / L}, j < L (the real transform has length 2L)
  int init(int L, bool half_angles = false);
  void release();
};

template <class R>
struct Fft3d {
  int N[3] = {0, 0, 0};
  int n[3] = {0, 0, 0};
  bool c2r = false;  // real grid: axis 2 is a real transform of length n2 (complex length n2 / 2)
  FftAxis<R> ax[3];
  int init(const int* N, const int* n, bool c2r = false);
  void release();
  // element counts of the two intermediate buffers (complex); c2r plans hold the
  // half spectrum H = N2/2 + 1 along axis 2 in both
  size_t buf1_elems() const {
    return c2r ? (size_t)N[0] * n[1] * (N[2] / 2 + 1) : (size_t)N[0] * N[1] * n[2];
  }
  size_t buf2_elems() const {
    return c2r ? (size_t)n[0] * n[1] * (N[2] / 2 + 1) : (size_t)N[0] * n[1] * n[2];
  }
  // complex elements of the grid's storage: n0 n1 n2 complex, or n0 n1 n2 reals (c2r)
  size_t grid_elems() const {
    const size_t nt = (size_t)n[0] * n[1] * n[2];
    return c2r ? (nt + 1) / 2 : nt;
  }
};

// grid <- F D fhat  (tab = concatenated FP64 D factors, or null to skip D).
// mark1 / mark2 (optional) are recorded on s after the first and second pass.
// half (c2r plans, F initialized with c2r): fhat holds the half spectrum k2 in
// [-N2/2, 0] and the grid is real, n0 n1 n2 reals in the grid's storage: the c2r
// transform (the k2 < 0 coefficients count twice, the k2 = 0 plane once).  Passes:
// axis 1 over the half spectrum (D fused), axis 0, then the c2r rows of axis 2.
// ctab (optional, overrides tab): complex per-axis factors, [N0 + N1 + N2] (re, im) FP64
// pairs in f_hat's index order; the value at (i0, i1, i2) is multiplied by their product.
template <class R>
int fft3d_forward(const Fft3d<R>& F, const Cplx<R>* fhat, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                  Cplx<R>* grid, hipStream_t s, bool half = false, hipEvent_t mark1 = nullptr,
                  hipEvent_t mark2 = nullptr, const double* ctab = nullptr);

// fhat (+)= D^T F^H grid; half: the grid is real and only k2 in [-N2/2, 0] is
// produced (r2c rows of axis 2, then axes 0 and 1, D^T fused into the last).
// ctab as for fft3d_forward, applied after the transform.
template <class R>
int fft3d_backward(const Fft3d<R>& F, const Cplx<R>* grid, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                   Cplx<R>* fhat, int accumulate, hipStream_t s, bool half = false, hipEvent_t mark1 = nullptr,
                   hipEvent_t mark2 = nullptr, const double* ctab = nullptr);

}  // namespace pnfft_amd
