// ndft.hip -- direct NDFT on the GPU (PNFFT_COMPUTE_DIRECT), the O(M N^3) sums
// of PNX(trafo_A) / PNX(adj_A) (kernel/ndft-parallel.c:377-722), c2c.
// Used for small problems and as an in-library cross-check; the exponentials
// e^{-2 pi i k_t x_t} are formed per axis from sincos in the compute precision.
#include "device_common.hpp"
#include "ndft.hpp"

namespace pnfft_amd {

namespace {

__device__ inline void sc(double a, double* s, double* c) { sincos(a, s, c); }
__device__ inline void sc(float a, float* s, float* c) { sincosf(a, s, c); }

template <class R>
__global__ __launch_bounds__(256) void ndft_trafo_k(NdftArgs<R> A) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= A.M) return;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2];
  const R twopi = R(6.28318530717958647692);
  const R x0 = A.x[3 * j], x1 = A.x[3 * j + 1], x2 = A.x[3 * j + 2];
  Cplx<R> f = cmk<R>(0, 0), g0 = cmk<R>(0, 0), g1 = cmk<R>(0, 0), g2 = cmk<R>(0, 0);
  for (int i0 = 0; i0 < N0; i0++) {
    const R k0 = R(i0 - N0 / 2);
    Cplx<R> s1 = cmk<R>(0, 0), s1k = cmk<R>(0, 0), s1z = cmk<R>(0, 0);
    for (int i1 = 0; i1 < N1; i1++) {
      const R k1 = R(i1 - N1 / 2);
      const Cplx<R>* fh = A.fhat + ((long)i0 * N1 + i1) * N2;
      Cplx<R> s2 = cmk<R>(0, 0), s2k = cmk<R>(0, 0);
      for (int i2 = 0; i2 < N2; i2++) {
        const R k2 = R(i2 - N2 / 2);
        R sn, cs;
        sc(-twopi * k2 * x2, &sn, &cs);
        const Cplx<R> p = cmul(fh[i2], cmk<R>(cs, sn));
        s2 = cadd(s2, p);
        s2k = cadd(s2k, cscale(p, k2));
      }
      R sn, cs;
      sc(-twopi * k1 * x1, &sn, &cs);
      const Cplx<R> e1 = cmk<R>(cs, sn);
      const Cplx<R> t = cmul(s2, e1);
      s1 = cadd(s1, t);
      s1k = cadd(s1k, cscale(t, k1));
      s1z = cadd(s1z, cmul(s2k, e1));
    }
    R sn, cs;
    sc(-twopi * k0 * x0, &sn, &cs);
    const Cplx<R> e0 = cmk<R>(cs, sn);
    const Cplx<R> t = cmul(s1, e0);
    f = cadd(f, t);
    g0 = cadd(g0, cscale(t, k0));
    g1 = cadd(g1, cmul(s1k, e0));
    g2 = cadd(g2, cmul(s1z, e0));
  }
  // the NDFT overwrites its outputs (kernel/ndft-parallel.c:400-414)
  if (A.want_f) A.f[j] = f;
  if (A.want_grad) {
    // grad *= -2 pi i (:594-603)
    A.grad[3 * j + 0] = cmk<R>(twopi * g0.y, -twopi * g0.x);
    A.grad[3 * j + 1] = cmk<R>(twopi * g1.y, -twopi * g1.x);
    A.grad[3 * j + 2] = cmk<R>(twopi * g2.y, -twopi * g2.x);
  }
}

// one thread per frequency; adds onto f_hat (:649-650, :681, :699-707)
template <class R>
__global__ __launch_bounds__(256) void ndft_adj_k(NdftArgs<R> A) {
  const long kidx = (long)blockIdx.x * 256 + threadIdx.x;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2];
  if (kidx >= (long)N0 * N1 * N2) return;
  const int i2 = (int)(kidx % N2), i1 = (int)((kidx / N2) % N1), i0 = (int)(kidx / ((long)N1 * N2));
  const R k0 = R(i0 - N0 / 2), k1 = R(i1 - N1 / 2), k2 = R(i2 - N2 / 2);
  const R twopi = R(6.28318530717958647692);
  Cplx<R> acc = cmk<R>(0, 0);
  for (long j = 0; j < A.M; j++) {
    R sn, cs;
    sc(twopi * (k0 * A.x[3 * j] + k1 * A.x[3 * j + 1] + k2 * A.x[3 * j + 2]), &sn, &cs);
    Cplx<R> c = cmk<R>(0, 0);
    if (A.want_f) c = A.f[j];
    if (A.want_grad) {
      const Cplx<R> g0 = A.grad[3 * j], g1 = A.grad[3 * j + 1], g2 = A.grad[3 * j + 2];
      const R sr = k0 * g0.x + k1 * g1.x + k2 * g2.x;
      const R si = k0 * g0.y + k1 * g1.y + k2 * g2.y;
      c = cadd(c, cmk<R>(-twopi * si, twopi * sr));
    }
    acc = cadd(acc, cmul(c, cmk<R>(cs, sn)));
  }
  A.fhat[kidx] = cadd(A.fhat[kidx], acc);
}

// is_hermitian of kernel/ndft-parallel.c:356-374: coefficients of the half
// spectrum that are self-conjugate or whose conjugate partner is also stored.
__device__ inline bool is_hermitian(int k0, int k1, int k2, int N0, int N1, int N2) {
  if ((k0 == 0 || k0 == -N0 / 2) && (k1 == 0 || k1 == -N1 / 2) && (k2 == 0 || k2 == -N2 / 2)) return true;
  if (k1 > 0 && (k2 == 0 || k2 == -N2 / 2)) return true;
  if (k0 > 0 && (k1 == 0 || k1 == -N1 / 2) && (k2 == 0 || k2 == -N2 / 2)) return true;
  return false;
}

// c2r trafo NDFT (:464-476, :524-533, :569-573, :594-603): f = Re f_0 + sum' 2 Re(f_k e^{-2 pi i k x}),
// grad_t = -2 pi sum' 2 k_t Im(f_k e^{-2 pi i k x}).
template <class R>
__global__ __launch_bounds__(256) void ndft_trafo_c2r_k(NdftArgs<R> A) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= A.M) return;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2], H2 = N2 / 2 + 1;
  const R twopi = R(6.28318530717958647692);
  const R x0 = A.x[3 * j], x1 = A.x[3 * j + 1], x2 = A.x[3 * j + 2];
  R f = 0, g0 = 0, g1 = 0, g2 = 0;
  for (int i0 = 0; i0 < N0; i0++)
    for (int i1 = 0; i1 < N1; i1++)
      for (int i2 = 0; i2 < H2; i2++) {
        const int k0 = i0 - N0 / 2, k1 = i1 - N1 / 2, k2 = i2 - N2 / 2;
        const Cplx<R> c = A.fhat[((long)i0 * N1 + i1) * H2 + i2];
        if (k0 == 0 && k1 == 0 && k2 == 0) {
          f += c.x;
        } else if (!is_hermitian(k0, k1, k2, N0, N1, N2)) {
          R sn, cs;
          sc(-twopi * (R(k0) * x0 + R(k1) * x1 + R(k2) * x2), &sn, &cs);
          const Cplx<R> z = cmul(c, cmk<R>(cs, sn));
          f += R(2) * z.x;
          g0 += R(2 * k0) * z.y;
          g1 += R(2 * k1) * z.y;
          g2 += R(2 * k2) * z.y;
        }
      }
  R* fr = reinterpret_cast<R*>(A.f);
  R* gr = reinterpret_cast<R*>(A.grad);
  if (A.want_f) fr[j] = f;
  if (A.want_grad) {
    gr[3 * j + 0] = -twopi * g0;
    gr[3 * j + 1] = -twopi * g1;
    gr[3 * j + 2] = -twopi * g2;
  }
}

// c2r adjoint NDFT: the c2c sums over the stored half with real f and grad (:667-715)
template <class R>
__global__ __launch_bounds__(256) void ndft_adj_c2r_k(NdftArgs<R> A) {
  const long kidx = (long)blockIdx.x * 256 + threadIdx.x;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2], H2 = N2 / 2 + 1;
  if (kidx >= (long)N0 * N1 * H2) return;
  const int i2 = (int)(kidx % H2), i1 = (int)((kidx / H2) % N1), i0 = (int)(kidx / ((long)N1 * H2));
  const R k0 = R(i0 - N0 / 2), k1 = R(i1 - N1 / 2), k2 = R(i2 - N2 / 2);
  const R twopi = R(6.28318530717958647692);
  const R* fr = reinterpret_cast<const R*>(A.f);
  const R* gr = reinterpret_cast<const R*>(A.grad);
  Cplx<R> acc = cmk<R>(0, 0);
  for (long j = 0; j < A.M; j++) {
    R sn, cs;
    sc(twopi * (k0 * A.x[3 * j] + k1 * A.x[3 * j + 1] + k2 * A.x[3 * j + 2]), &sn, &cs);
    Cplx<R> c = cmk<R>(A.want_f ? fr[j] : R(0), R(0));
    if (A.want_grad) c.y += twopi * (k0 * gr[3 * j] + k1 * gr[3 * j + 1] + k2 * gr[3 * j + 2]);
    acc = cadd(acc, cmul(c, cmk<R>(cs, sn)));
  }
  A.fhat[kidx] = cadd(A.fhat[kidx], acc);
}

// Hessian NDFT (:450-510, scaled by -4 pi^2 at :604-612): hess[6j + q] =
// -4 pi^2 sum_k k_a k_b f_hat_k e^{-2 pi i k x_j}, (a, b) = xx, xy, xz, yy, yz, zz;
// c2c nested per axis as ndft_trafo_k, c2r over the half spectrum with weight 2
// and the skips of is_hermitian.  Overwrites hess.
template <class R>
__global__ __launch_bounds__(256) void ndft_hess_k(NdftArgs<R> A) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= A.M) return;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2];
  const R twopi = R(6.28318530717958647692);
  const R x0 = A.x[3 * j], x1 = A.x[3 * j + 1], x2 = A.x[3 * j + 2];
  // level sums: s2[e] = sum_i2 k2^e p, then per i1 and i0 the products with k1, k0
  Cplx<R> h[6];
  for (int q = 0; q < 6; q++) h[q] = cmk<R>(0, 0);
  for (int i0 = 0; i0 < N0; i0++) {
    const R k0 = R(i0 - N0 / 2);
    // t[0] = sum e1 s2_0, t[1] = sum k1 e1 s2_0, t[2] = sum k1^2 e1 s2_0,
    // t[3] = sum e1 s2_1, t[4] = sum k1 e1 s2_1, t[5] = sum e1 s2_2
    Cplx<R> t[6];
    for (int q = 0; q < 6; q++) t[q] = cmk<R>(0, 0);
    for (int i1 = 0; i1 < N1; i1++) {
      const R k1 = R(i1 - N1 / 2);
      const Cplx<R>* fh = A.fhat + ((long)i0 * N1 + i1) * N2;
      Cplx<R> u0 = cmk<R>(0, 0), u1 = cmk<R>(0, 0), u2 = cmk<R>(0, 0);
      for (int i2 = 0; i2 < N2; i2++) {
        const R k2 = R(i2 - N2 / 2);
        R sn, cs;
        sc(-twopi * k2 * x2, &sn, &cs);
        const Cplx<R> pv = cmul(fh[i2], cmk<R>(cs, sn));
        u0 = cadd(u0, pv);
        u1 = cadd(u1, cscale(pv, k2));
        u2 = cadd(u2, cscale(pv, k2 * k2));
      }
      R sn, cs;
      sc(-twopi * k1 * x1, &sn, &cs);
      const Cplx<R> e1 = cmk<R>(cs, sn);
      const Cplx<R> v0 = cmul(u0, e1), v1 = cmul(u1, e1), v2 = cmul(u2, e1);
      t[0] = cadd(t[0], v0);
      t[1] = cadd(t[1], cscale(v0, k1));
      t[2] = cadd(t[2], cscale(v0, k1 * k1));
      t[3] = cadd(t[3], v1);
      t[4] = cadd(t[4], cscale(v1, k1));
      t[5] = cadd(t[5], v2);
    }
    R sn, cs;
    sc(-twopi * k0 * x0, &sn, &cs);
    const Cplx<R> e0 = cmk<R>(cs, sn);
    h[0] = cadd(h[0], cscale(cmul(t[0], e0), k0 * k0));  // xx
    h[1] = cadd(h[1], cscale(cmul(t[1], e0), k0));       // xy
    h[2] = cadd(h[2], cscale(cmul(t[3], e0), k0));       // xz
    h[3] = cadd(h[3], cmul(t[2], e0));                   // yy
    h[4] = cadd(h[4], cmul(t[4], e0));                   // yz
    h[5] = cadd(h[5], cmul(t[5], e0));                   // zz
  }
  const R c = -twopi * twopi;
  for (int q = 0; q < 6; q++) A.hess[6 * j + q] = cscale(h[q], c);
}

template <class R>
__global__ __launch_bounds__(256) void ndft_hess_c2r_k(NdftArgs<R> A) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= A.M) return;
  const int N0 = A.N[0], N1 = A.N[1], N2 = A.N[2], H2 = N2 / 2 + 1;
  const R twopi = R(6.28318530717958647692);
  const R x0 = A.x[3 * j], x1 = A.x[3 * j + 1], x2 = A.x[3 * j + 2];
  R h[6] = {0, 0, 0, 0, 0, 0};
  for (int i0 = 0; i0 < N0; i0++)
    for (int i1 = 0; i1 < N1; i1++)
      for (int i2 = 0; i2 < H2; i2++) {
        const int k0 = i0 - N0 / 2, k1 = i1 - N1 / 2, k2 = i2 - N2 / 2;
        if ((k0 == 0 && k1 == 0 && k2 == 0) || is_hermitian(k0, k1, k2, N0, N1, N2)) continue;
        const Cplx<R> cf = A.fhat[((long)i0 * N1 + i1) * H2 + i2];
        R sn, cs;
        sc(-twopi * (R(k0) * x0 + R(k1) * x1 + R(k2) * x2), &sn, &cs);
        const R re = R(2) * cmul(cf, cmk<R>(cs, sn)).x;
        h[0] += R(k0 * k0) * re;
        h[1] += R(k0 * k1) * re;
        h[2] += R(k0 * k2) * re;
        h[3] += R(k1 * k1) * re;
        h[4] += R(k1 * k2) * re;
        h[5] += R(k2 * k2) * re;
      }
  R* hr = reinterpret_cast<R*>(A.hess);
  for (int q = 0; q < 6; q++) hr[6 * j + q] = -twopi * twopi * h[q];
}

}  // namespace

template <class R>
int ndft_trafo(const NdftArgs<R>& A, hipStream_t s) {
  if (A.M == 0) return kOk;
  const dim3 grid((unsigned)((A.M + 255) / 256)), block(256);
  if (A.want_f || A.want_grad) {
    if (A.c2r)
      hipLaunchKernelGGL((ndft_trafo_c2r_k<R>), grid, block, 0, s, A);
    else
      hipLaunchKernelGGL((ndft_trafo_k<R>), grid, block, 0, s, A);
  }
  if (A.want_hess) {
    if (A.c2r)
      hipLaunchKernelGGL((ndft_hess_c2r_k<R>), grid, block, 0, s, A);
    else
      hipLaunchKernelGGL((ndft_hess_k<R>), grid, block, 0, s, A);
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int ndft_adj(const NdftArgs<R>& A, hipStream_t s) {
  const long K = (long)A.N[0] * A.N[1] * (A.c2r ? A.N[2] / 2 + 1 : A.N[2]);
  if (K == 0) return kOk;
  if (A.c2r)
    hipLaunchKernelGGL((ndft_adj_c2r_k<R>), dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL((ndft_adj_k<R>), dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, A);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template int ndft_trafo<float>(const NdftArgs<float>&, hipStream_t);
template int ndft_trafo<double>(const NdftArgs<double>&, hipStream_t);
template int ndft_adj<float>(const NdftArgs<float>&, hipStream_t);
template int ndft_adj<double>(const NdftArgs<double>&, hipStream_t);

}  // namespace pnfft_amd
