// b_kernels.hip -- gather (trafo B) and scatter (adj B^T) kernels.
//
// This file is compiled once per (precision, kind, mode); the build passes
// PNFFT_B_REAL (float|double), PNFFT_B_KIND (0 gather, 1 scatter) and
// PNFFT_B_MODE (BMode), and the file instantiates the launcher
// b_launch<PNFFT_B_REAL, PNFFT_B_KIND, PNFFT_B_MODE> for them.
//
// Reference: the node loops of kernel/ndft-parallel.c:2703-3009 and the
// stencil kernels of kernel/assign.c:478-1066 (cutoff 2m+1, :910).
#include "b_common.hpp"
#include "device_common.hpp"
#include "nfft_b.hpp"
#include "window_math.hpp"

namespace pnfft_amd {

namespace {

constexpr int kB = 256;  // threads per block

// ---------------------------------------------------------------------------
// direct gather: one thread per node, stencil read from the grid in HBM/L2
// ---------------------------------------------------------------------------
template <class R, int MT, int MODE>
__global__ __launch_bounds__(kB) void gather_direct_k(BArgs<R> A) {
  const long p = (long)blockIdx.x * kB + threadIdx.x;
  if (p >= A.M) return;
  const long j = A.perm ? (long)A.perm[p] : p;
  const int m = MT > 0 ? MT : A.g.m;
  constexpr int CM = MT > 0 ? 2 * MT + 1 : kMaxC;
  const int c = 2 * m + 1;
  constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD || MODE == B_VJP);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];

  R px[CM], py[CM], pz[CM], dx[CM], dy[CM], dz[CM];
  const NodeAxis<R> a0 = node_axis<R>(A.x[3 * j + 0], n0, m);
  const NodeAxis<R> a1 = node_axis<R>(A.x[3 * j + 1], n1, m);
  const NodeAxis<R> a2 = node_axis<R>(A.x[3 * j + 2], n2, m);
  window_axis<R>(A.g.win, m, a0.frac, (R)n0, (R)A.g.b[0], px, WG ? dx : nullptr);
  window_axis<R>(A.g.win, m, a1.frac, (R)n1, (R)A.g.b[1], py, WG ? dy : nullptr);
  window_axis<R>(A.g.win, m, a2.frac, (R)n2, (R)A.g.b[2], pz, WG ? dz : nullptr);

  int i2s[CM];
#pragma unroll
  for (int s = 0; s < CM; s++) {
    if (s < c) {
      int v = a2.start + s;
      i2s[s] = v >= n2 ? v - n2 : v;
    }
  }

  Cplx<R> f = cmk<R>(0, 0), g0 = cmk<R>(0, 0), g1 = cmk<R>(0, 0), g2 = cmk<R>(0, 0);
  for (int a = 0; a < c; a++) {
    int i0 = a0.start + a;
    if (i0 >= n0) i0 -= n0;
    Cplx<R> ha = cmk<R>(0, 0), hay = cmk<R>(0, 0), haz = cmk<R>(0, 0);
    for (int b = 0; b < c; b++) {
      int i1 = a1.start + b;
      if (i1 >= n1) i1 -= n1;
      const Cplx<R>* row = A.grid + ((long)i0 * n1 + i1) * n2;
      Cplx<R> hb = cmk<R>(0, 0), hbz = cmk<R>(0, 0);
#pragma unroll
      for (int s = 0; s < CM; s++) {
        if (s < c) {
          const Cplx<R> gv = row[i2s[s]];
          hb.x += pz[s] * gv.x;
          hb.y += pz[s] * gv.y;
          if (WG) {
            hbz.x += dz[s] * gv.x;
            hbz.y += dz[s] * gv.y;
          }
        }
      }
      ha.x += py[b] * hb.x;
      ha.y += py[b] * hb.y;
      if (WG) {
        hay.x += dy[b] * hb.x;
        hay.y += dy[b] * hb.y;
        haz.x += py[b] * hbz.x;
        haz.y += py[b] * hbz.y;
      }
    }
    f.x += px[a] * ha.x;
    f.y += px[a] * ha.y;
    if (WG) {
      g0.x += dx[a] * ha.x;
      g0.y += dx[a] * ha.y;
      g1.x += px[a] * hay.x;
      g1.y += px[a] * hay.y;
      g2.x += px[a] * haz.x;
      g2.y += px[a] * haz.y;
    }
  }
  if (WF) {
    Cplx<R> o = A.accumulate ? A.f[j] : cmk<R>(0, 0);
    if (A.real_only) f.y = 0;  // assign_f_r2r reads the real grid part only (kernel/assign.c:715-739)
    A.f[j] = cadd(o, f);
  }
  if (MODE == B_GRAD || MODE == B_FGRAD) {
    Cplx<R>* gp = A.grad + 3 * j;
    if (A.accumulate) {
      g0 = cadd(g0, gp[0]);
      g1 = cadd(g1, gp[1]);
      g2 = cadd(g2, gp[2]);
    }
    gp[0] = g0;
    gp[1] = g1;
    gp[2] = g2;
  }
  if (MODE == B_VJP) {
    const Cplx<R> w = A.w[j];
    A.gx[3 * j + 0] = w.x * g0.x + w.y * g0.y;
    A.gx[3 * j + 1] = w.x * g1.x + w.y * g1.y;
    A.gx[3 * j + 2] = w.x * g2.x + w.y * g2.y;
  }
}

// ---------------------------------------------------------------------------
// direct scatter: one thread per node, hardware float atomics into the grid
// ---------------------------------------------------------------------------
template <class R, int MT, int MODE>
__global__ __launch_bounds__(kB) void scatter_direct_k(BArgs<R> A) {
  const long p = (long)blockIdx.x * kB + threadIdx.x;
  if (p >= A.M) return;
  const long j = A.perm ? (long)A.perm[p] : p;
  const int m = MT > 0 ? MT : A.g.m;
  constexpr int CM = MT > 0 ? 2 * MT + 1 : kMaxC;
  const int c = 2 * m + 1;
  constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];

  R px[CM], py[CM], pz[CM], dx[CM], dy[CM], dz[CM];
  const NodeAxis<R> a0 = node_axis<R>(A.x[3 * j + 0], n0, m);
  const NodeAxis<R> a1 = node_axis<R>(A.x[3 * j + 1], n1, m);
  const NodeAxis<R> a2 = node_axis<R>(A.x[3 * j + 2], n2, m);
  window_axis<R>(A.g.win, m, a0.frac, (R)n0, (R)A.g.b[0], px, WG ? dx : nullptr);
  window_axis<R>(A.g.win, m, a1.frac, (R)n1, (R)A.g.b[1], py, WG ? dy : nullptr);
  window_axis<R>(A.g.win, m, a2.frac, (R)n2, (R)A.g.b[2], pz, WG ? dz : nullptr);

  Cplx<R> fv = cmk<R>(0, 0), gv0 = cmk<R>(0, 0), gv1 = cmk<R>(0, 0), gv2 = cmk<R>(0, 0);
  if (WF) fv = A.f[j];
  if (WG) {
    gv0 = A.grad[3 * j + 0];
    gv1 = A.grad[3 * j + 1];
    gv2 = A.grad[3 * j + 2];
  }
  for (int a = 0; a < c; a++) {
    int i0 = a0.start + a;
    if (i0 >= n0) i0 -= n0;
    for (int b = 0; b < c; b++) {
      int i1 = a1.start + b;
      if (i1 >= n1) i1 -= n1;
      Cplx<R>* row = A.grid + ((long)i0 * n1 + i1) * n2;
      const R pxy = px[a] * py[b];
      R pdxy = 0, pxdy = 0;
      if (WG) {
        pdxy = dx[a] * py[b];
        pxdy = px[a] * dy[b];
      }
      for (int s = 0; s < c; s++) {
        int i2 = a2.start + s;
        if (i2 >= n2) i2 -= n2;
        R vr = 0, vi = 0;
        if (WF) {
          const R w = pxy * pz[s];
          vr += w * fv.x;
          vi += w * fv.y;
        }
        if (WG) {
          const R w0 = pdxy * pz[s], w1 = pxdy * pz[s], w2 = pxy * dz[s];
          vr += w0 * gv0.x + w1 * gv1.x + w2 * gv2.x;
          vi += w0 * gv0.y + w1 * gv1.y + w2 * gv2.y;
        }
        atomicAdd(&row[i2].x, vr);
        atomicAdd(&row[i2].y, vi);
      }
    }
  }
}

inline unsigned grid1d(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

template <class R, int KIND, int MODE>
int b_launch(const BArgs<R>& A, hipStream_t s) {
  const dim3 grid(grid1d(A.M, kB)), block(kB);
  switch (A.g.m) {
#define PNFFT_B_CASE(MV)                                                                       \
  case MV:                                                                                     \
    if (KIND == 0) hipLaunchKernelGGL((gather_direct_k<R, MV, MODE>), grid, block, 0, s, A);   \
    else hipLaunchKernelGGL((scatter_direct_k<R, MV, (MODE > 3 ? 3 : MODE)>), grid, block, 0, s, A); \
    break;
    PNFFT_B_CASE(2)
    PNFFT_B_CASE(3)
    PNFFT_B_CASE(4)
    PNFFT_B_CASE(5)
    PNFFT_B_CASE(6)
    PNFFT_B_CASE(7)
    PNFFT_B_CASE(8)
#undef PNFFT_B_CASE
    default:
      if (KIND == 0) hipLaunchKernelGGL((gather_direct_k<R, 0, MODE>), grid, block, 0, s, A);
      else hipLaunchKernelGGL((scatter_direct_k<R, 0, (MODE > 3 ? 3 : MODE)>), grid, block, 0, s, A);
      break;
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template int b_launch<PNFFT_B_REAL, PNFFT_B_KIND, PNFFT_B_MODE>(const BArgs<PNFFT_B_REAL>&, hipStream_t);

}  // namespace pnfft_amd
