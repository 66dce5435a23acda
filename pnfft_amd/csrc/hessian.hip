// hessian.hip -- the Hessian gather of PNFFT_COMPUTE_HESSIAN_F and the strided
// node-value copies of the ik-differentiated transforms.
//
// This file is compiled once per precision; the build passes PNFFT_B_REAL.
//
// Reference: loop_over_particles_trafo (kernel/ndft-parallel.c:2703-2888) with
// assign_hessian_f_c2c_pre_psi / _r2r (kernel/assign.c:881-1027) and the
// second-derivative windows pre_ddpsi_tensor_* (kernel/ndft-parallel.c:1956-2105).
// hess[6j + q], q = xx, xy, xz, yy, yz, zz, is
//   sum_{a,b,c} X_q(a) Y_q(b) Z_q(c) g[a][b][c]
// with (X, Y, Z) = (psi'', psi, psi), (psi', psi', psi), (psi', psi, psi'),
// (psi, psi'', psi), (psi, psi', psi'), (psi, psi, psi'').
//
// Two kernels: the first evaluates psi, psi' and psi'' of the three axes of
// every node (in the sorted order of the node bins) into a table, structure of
// arrays, so the second one, a gather straight from the grid in HBM (periodic
// wrap, one thread per node), keeps only the z-axis values in registers and
// reads the x and y factors from the table as its plane and row loops reach
// them.  This is not the headline path; it is written for register bounds that
// hold at every m = 2..8 in both precisions rather than for speed.
#include "b_device.hpp"
#include "window_ext.hpp"

namespace pnfft_amd {

namespace {

using namespace bdev;

constexpr int kHT = 256;  // threads per block

// tab[((t * 3 + q) * CM + s) * M + p], q = 0 psi, 1 psi', 2 psi''; st[t * M + p]
template <class R, int MT>
__global__ __launch_bounds__(kHT) void hess_tab_k(BArgs<R> A, IntpolTabs<R> I, R* __restrict__ tab,
                                                  int* __restrict__ st) {
  constexpr int CM = 2 * MT + 1;
  const long p = (long)blockIdx.x * kHT + threadIdx.x;
  const long M = A.M;
  if (p >= M) return;
  const long j = A.perm ? (long)A.perm[p] : p;
#pragma unroll 1
  for (int t = 0; t < 3; t++) {
    const int nt = A.g.n[t];
    const NodeAxis<R> a = node_axis<R>(A.x[3 * j + t], nt, MT, A.g.shift);
    R psi[CM], dpsi[CM], dd[CM];
    ext_window_axis<R, CM, true, true>(A.g, I, t, a.frac, psi, dpsi, dd);
    const R ws = (R)A.g.ws[t];  // the plan's power-of-two window scale (exact)
#pragma unroll
    for (int s = 0; s < CM; s++) {
      tab[((long)(t * 3 + 0) * CM + s) * M + p] = psi[s] * ws;
      tab[((long)(t * 3 + 1) * CM + s) * M + p] = dpsi[s] * ws;
      tab[((long)(t * 3 + 2) * CM + s) * M + p] = dd[s] * ws;
    }
    st[(long)t * M + p] = a.start;
  }
}

template <class R, int MT>
__global__ __launch_bounds__(kHT) void hess_gather_k(BArgs<R> A, Cplx<R>* __restrict__ hess,
                                                     const R* __restrict__ tab, const int* __restrict__ st) {
  constexpr int CM = 2 * MT + 1;
  const long p = (long)blockIdx.x * kHT + threadIdx.x;
  const long M = A.M;
  if (p >= M) return;
  const long j = A.perm ? (long)A.perm[p] : p;
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  auto T = [&](int t, int q, int s) { return tab[((long)(t * 3 + q) * CM + s) * M + p]; };
  R z0[CM], z1[CM], z2[CM];
  int i2s[CM];
  const int s0 = st[p], s1 = st[M + p], s2 = st[2 * M + p];
#pragma unroll
  for (int s = 0; s < CM; s++) {
    z0[s] = T(2, 0, s);
    z1[s] = T(2, 1, s);
    z2[s] = T(2, 2, s);
    const int v = s2 + s;
    i2s[s] = v >= n2 ? v - n2 : v;
  }
  // h[q] accumulates hessian component q (re, im)
  R hr[6] = {0, 0, 0, 0, 0, 0}, hi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
  for (int a = 0; a < CM; a++) {
    int i0 = s0 + a;
    if (i0 >= n0) i0 -= n0;
    const R x0 = T(0, 0, a), x1 = T(0, 1, a), x2 = T(0, 2, a);
    R ur[6] = {0, 0, 0, 0, 0, 0}, ui[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int b = 0; b < CM; b++) {
      int i1 = s1 + b;
      if (i1 >= n1) i1 -= n1;
      const long row = ((long)i0 * n1 + i1) * n2;
      R a0r = 0, a0i = 0, a1r = 0, a1i = 0, a2r = 0, a2i = 0;  // sum_c z G, z' G, z'' G
#pragma unroll
      for (int s = 0; s < CM; s++) {
        // c2r plans: the grid is real
        const Cplx<R> gv = A.c2r ? cmk<R>(reinterpret_cast<const R*>(A.grid)[row + i2s[s]], R(0))
                                 : A.grid[row + i2s[s]];
        const R gr = gv.x, gi = gv.y;
        a0r += z0[s] * gr;
        a0i += z0[s] * gi;
        a1r += z1[s] * gr;
        a1i += z1[s] * gi;
        a2r += z2[s] * gr;
        a2i += z2[s] * gi;
      }
      // U[q] = sum_b Y_q(b) (sum_c Z_q(c) g): the y and z factors of component q
      const R y0 = T(1, 0, b), y1 = T(1, 1, b), y2 = T(1, 2, b);
      ur[0] += y0 * a0r;  // xx: psi_y psi_z
      ui[0] += y0 * a0i;
      ur[1] += y1 * a0r;  // xy: psi'_y psi_z
      ui[1] += y1 * a0i;
      ur[2] += y0 * a1r;  // xz: psi_y psi'_z
      ui[2] += y0 * a1i;
      ur[3] += y2 * a0r;  // yy: psi''_y psi_z
      ui[3] += y2 * a0i;
      ur[4] += y1 * a1r;  // yz: psi'_y psi'_z
      ui[4] += y1 * a1i;
      ur[5] += y0 * a2r;  // zz: psi_y psi''_z
      ui[5] += y0 * a2i;
    }
    hr[0] += x2 * ur[0];  // xx = psi''_x psi_y psi_z
    hi[0] += x2 * ui[0];
    hr[1] += x1 * ur[1];  // xy = psi'_x psi'_y psi_z
    hi[1] += x1 * ui[1];
    hr[2] += x1 * ur[2];  // xz = psi'_x psi_y psi'_z
    hi[2] += x1 * ui[2];
    hr[3] += x0 * ur[3];  // yy = psi_x psi''_y psi_z
    hi[3] += x0 * ui[3];
    hr[4] += x0 * ur[4];  // yz = psi_x psi'_y psi'_z
    hi[4] += x0 * ui[4];
    hr[5] += x0 * ur[5];  // zz = psi_x psi_y psi''_z
    hi[5] += x0 * ui[5];
  }
  if (A.c2r) {  // real plan: R[6M] (assign_hessian_f_r2r)
    R* h = reinterpret_cast<R*>(hess) + 6 * j;
#pragma unroll
    for (int q = 0; q < 6; q++) h[q] = (A.accumulate ? h[q] : R(0)) + hr[q];
    return;
  }
  Cplx<R>* h = hess + 6 * j;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const Cplx<R> o = A.accumulate ? h[q] : cmk<R>(0, 0);
    // PNFFT_REAL_F: the real part only, the imaginary part is left as it was
    h[q] = A.real_only ? cmk<R>(o.x + hr[q], o.y) : cmk<R>(o.x + hr[q], o.y + hi[q]);
  }
}

template <class R, int MT>
int launch_hess(const BArgs<R>& A, const IntpolTabs<R>& I, Cplx<R>* hess, R* tab, int* st, hipStream_t s) {
  const dim3 grid((unsigned)((A.M + kHT - 1) / kHT)), block(kHT);
  hipLaunchKernelGGL((hess_tab_k<R, MT>), grid, block, 0, s, A, I, tab, st);
  hipLaunchKernelGGL((hess_gather_k<R, MT>), grid, block, 0, s, A, hess, tab, st);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

// out[(j stride + offset) comps + c] (+)= in[j comps + c]
template <class R>
__global__ __launch_bounds__(256) void strided_put_k(R* __restrict__ out, const R* __restrict__ in, long M,
                                                     int comps, int stride, int offset, int acc) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= M) return;
  for (int c = 0; c < comps; c++) {
    const long o = (j * stride + offset) * comps + c;
    out[o] = (acc ? out[o] : R(0)) + in[j * comps + c];
  }
}

// out[j comps + c] = in[(j stride + offset) comps + c]
template <class R>
__global__ __launch_bounds__(256) void strided_get_k(R* __restrict__ out, const R* __restrict__ in, long M,
                                                     int comps, int stride, int offset) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= M) return;
  for (int c = 0; c < comps; c++) out[j * comps + c] = in[(j * stride + offset) * comps + c];
}

}  // namespace

template <class R>
int hessian_gather(const BArgs<R>& A, const IntpolTabs<R>& I, Cplx<R>* hess, R* tab, int* st, hipStream_t s) {
  if (A.M == 0) return kOk;
  switch (A.g.m) {
    case 2: return launch_hess<R, 2>(A, I, hess, tab, st, s);
    case 3: return launch_hess<R, 3>(A, I, hess, tab, st, s);
    case 4: return launch_hess<R, 4>(A, I, hess, tab, st, s);
    case 5: return launch_hess<R, 5>(A, I, hess, tab, st, s);
    case 6: return launch_hess<R, 6>(A, I, hess, tab, st, s);
    case 7: return launch_hess<R, 7>(A, I, hess, tab, st, s);
    case 8: return launch_hess<R, 8>(A, I, hess, tab, st, s);
    default: return kUnsupported;
  }
}

template <class R>
int strided_put(R* out, const R* in, long M, int comps, int stride, int offset, int acc, hipStream_t s) {
  if (M == 0) return kOk;
  hipLaunchKernelGGL((strided_put_k<R>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, out, in, M, comps,
                     stride, offset, acc);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int strided_get(R* out, const R* in, long M, int comps, int stride, int offset, hipStream_t s) {
  if (M == 0) return kOk;
  hipLaunchKernelGGL((strided_get_k<R>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, out, in, M, comps,
                     stride, offset);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template int hessian_gather<PNFFT_B_REAL>(const BArgs<PNFFT_B_REAL>&, const IntpolTabs<PNFFT_B_REAL>&,
                                          Cplx<PNFFT_B_REAL>*, PNFFT_B_REAL*, int*, hipStream_t);
template int strided_put<PNFFT_B_REAL>(PNFFT_B_REAL*, const PNFFT_B_REAL*, long, int, int, int, int, hipStream_t);
template int strided_get<PNFFT_B_REAL>(PNFFT_B_REAL*, const PNFFT_B_REAL*, long, int, int, int, hipStream_t);

}  // namespace pnfft_amd
