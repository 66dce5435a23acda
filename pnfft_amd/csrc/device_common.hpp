// device_common.hpp -- types shared by the HIP kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pnfft_amd {

template <class R> struct Cplx;
template <> struct alignas(8) Cplx<float> { float x, y; };
template <> struct alignas(16) Cplx<double> { double x, y; };

template <class R>
__host__ __device__ inline Cplx<R> cmk(R a, R b) {
  Cplx<R> r;
  r.x = a;
  r.y = b;
  return r;
}
template <class R>
__host__ __device__ inline Cplx<R> cadd(Cplx<R> a, Cplx<R> b) { return cmk<R>(a.x + b.x, a.y + b.y); }
template <class R>
__host__ __device__ inline Cplx<R> csub(Cplx<R> a, Cplx<R> b) { return cmk<R>(a.x - b.x, a.y - b.y); }
template <class R>
__host__ __device__ inline Cplx<R> cmul(Cplx<R> a, Cplx<R> b) {
  return cmk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class R>
__host__ __device__ inline Cplx<R> cconj(Cplx<R> a) { return cmk<R>(a.x, -a.y); }
template <class R>
__host__ __device__ inline Cplx<R> cscale(Cplx<R> a, R s) { return cmk<R>(a.x * s, a.y * s); }

// Geometry of one oversampled grid and its window, passed to the B kernels.
struct GridGeom {
  int n[3];      // oversampled grid (natural DFT order: index l mod n)
  int m;         // window cutoff parameter (stencil 2m+1)
  int win;       // WindowId
  double b[3];   // window shape parameter per axis
  // power-of-two factor applied to the window values of each axis in the B
  // kernels (2^-e with e the exponent of the window's peak) and undone in the
  // D factors, so that FP32 intermediates stay near 1; exact in both precisions
  double ws[3];
  // added to n_t x_t before the stencil start and the window argument are taken:
  // 0, or 1/2 for the second pass of an interlaced plan (nodes shifted by half a
  // cell, kernel/ndft-parallel.c:2732-2753)
  double shift;
};

// Launch helpers report errors through this code (0 = ok).
enum Status : int { kOk = 0, kUnsupported = 1, kHipError = 2 };

}  // namespace pnfft_amd
