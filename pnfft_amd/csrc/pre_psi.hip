// pre_psi.hip -- window tables of pnfft_precompute_psi.
//
// This file is compiled once per precision; the build passes PNFFT_B_REAL.
//
// Reference: PNX(precompute_psi) (kernel/ndft-parallel.c:1144-1245) fills, per
// node and in the order matrix B visits the nodes, the tensor-product factors
// psi_t(s), t = 0..2, s = 0..2m (3 (2m+1) values, :1345-1354).  Here the tables
// are stored structure-of-arrays in the sorted order of the node bins,
// psi[(t CM + s) M + p] for sorted position p, so a wavefront reads them with
// coalesced loads; the coordinates they were computed from are kept beside them
// (x_out[t M + p]) so that the B kernels can tell whether a node moved since.
// With a dpsi array the derivative factors are tabulated the same way (the
// reference allocates them for PNFFT_PRE_GRAD_PSI, :1186, but only under
// PNFFT_DIFF_AD, :1153-1156).
#include "b_device.hpp"

namespace pnfft_amd {

namespace {

using namespace bdev;

template <class R, int MT, bool WD>
__global__ __launch_bounds__(256) void pre_psi_k(BArgs<R> A, R* __restrict__ psi, R* __restrict__ dpsi,
                                                 R* __restrict__ x_out) {
  constexpr int CM = 2 * MT + 1;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long M = A.M;
  if (p >= M) return;
  const long j = A.perm[p];
  NodeWin<R, CM> w;
  node_windows<R, MT, CM, WD>(A, -1, j, MT, w);
#pragma unroll
  for (int s = 0; s < CM; s++) {
    psi[(long)s * M + p] = w.px[s];
    psi[(long)(CM + s) * M + p] = w.py[s];
    psi[(long)(2 * CM + s) * M + p] = w.pz[s];
    if (WD) {
      dpsi[(long)s * M + p] = w.dx[s];
      dpsi[(long)(CM + s) * M + p] = w.dy[s];
      dpsi[(long)(2 * CM + s) * M + p] = w.dz[s];
    }
  }
#pragma unroll
  for (int t = 0; t < 3; t++) x_out[(long)t * M + p] = A.x[3 * j + t];
}

template <class R, int MT>
int launch_tables(const BArgs<R>& A, R* psi, R* dpsi, R* x_out, hipStream_t s) {
  if constexpr (!static_body<R, MT>()) {
    return kUnsupported;
  } else {
    const dim3 grid((unsigned)((A.M + 255) / 256)), block(256);
    if (dpsi)
      hipLaunchKernelGGL((pre_psi_k<R, MT, true>), grid, block, 0, s, A, psi, dpsi, x_out);
    else
      hipLaunchKernelGGL((pre_psi_k<R, MT, false>), grid, block, 0, s, A, psi, dpsi, x_out);
    return hipGetLastError() == hipSuccess ? kOk : kHipError;
  }
}

}  // namespace

template <class R>
int precompute_tables(const BArgs<R>& A, R* psi, R* dpsi, R* x_out, hipStream_t s) {
  if (A.M == 0 || !A.perm || A.pre_psi) return kUnsupported;
  switch (A.g.m) {
    case 2: return launch_tables<R, 2>(A, psi, dpsi, x_out, s);
    case 3: return launch_tables<R, 3>(A, psi, dpsi, x_out, s);
    case 4: return launch_tables<R, 4>(A, psi, dpsi, x_out, s);
    case 5: return launch_tables<R, 5>(A, psi, dpsi, x_out, s);
    case 6: return launch_tables<R, 6>(A, psi, dpsi, x_out, s);
    case 7: return launch_tables<R, 7>(A, psi, dpsi, x_out, s);
    case 8: return launch_tables<R, 8>(A, psi, dpsi, x_out, s);
    default: return kUnsupported;
  }
}

template int precompute_tables<PNFFT_B_REAL>(const BArgs<PNFFT_B_REAL>&, PNFFT_B_REAL*, PNFFT_B_REAL*,
                                             PNFFT_B_REAL*, hipStream_t);

}  // namespace pnfft_amd
