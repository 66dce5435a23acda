// window_tab.hip -- per-node window tables of the plans whose windows the B
// kernels do not evaluate themselves (the Bessel-I0 window, interpolated
// windows; window_ext.hpp).  Such plans tabulate psi (and psi') of every node
// in the bin order before each B stage, in the layout of pnfft_precompute_psi's
// tables, so the gather and scatter kernels run unchanged and read the tables.
//
// This file is compiled once per precision; the build passes PNFFT_B_REAL.
#include "b_device.hpp"
#include "window_ext.hpp"

namespace pnfft_amd {

namespace {

using namespace bdev;

template <class R, int MT, bool WD>
__global__ __launch_bounds__(256) void ext_tab_k(BArgs<R> A, IntpolTabs<R> I, R* __restrict__ psi,
                                                 R* __restrict__ dpsi, R* __restrict__ x_out) {
  constexpr int CM = 2 * MT + 1;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long M = A.M;
  if (p >= M) return;
  const long j = A.perm[p];
#pragma unroll 1
  for (int t = 0; t < 3; t++) {
    const NodeAxis<R> a = node_axis<R>(A.x[3 * j + t], A.g.n[t], MT, A.g.shift);
    R ps[CM], dp[CM], dd[CM];
    ext_window_axis<R, CM, WD, false>(A.g, I, t, a.frac, ps, dp, dd);
    const R ws = (R)A.g.ws[t];  // the plan's power-of-two window scale (exact)
#pragma unroll
    for (int s = 0; s < CM; s++) {
      psi[(long)(t * CM + s) * M + p] = ps[s] * ws;
      if (WD) dpsi[(long)(t * CM + s) * M + p] = dp[s] * ws;
    }
    x_out[(long)t * M + p] = A.x[3 * j + t];
  }
}

template <class R, int MT>
int launch_ext(const BArgs<R>& A, const IntpolTabs<R>& I, R* psi, R* dpsi, R* x_out, hipStream_t s) {
  const dim3 grid((unsigned)((A.M + 255) / 256)), block(256);
  if (dpsi)
    hipLaunchKernelGGL((ext_tab_k<R, MT, true>), grid, block, 0, s, A, I, psi, dpsi, x_out);
  else
    hipLaunchKernelGGL((ext_tab_k<R, MT, false>), grid, block, 0, s, A, I, psi, dpsi, x_out);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

}  // namespace

template <class R>
int ext_tables(const BArgs<R>& A, const IntpolTabs<R>& I, R* psi, R* dpsi, R* x_out, hipStream_t s) {
  if (A.M == 0) return kOk;
  if (!A.perm) return kUnsupported;
  switch (A.g.m) {
    case 2: return launch_ext<R, 2>(A, I, psi, dpsi, x_out, s);
    case 3: return launch_ext<R, 3>(A, I, psi, dpsi, x_out, s);
    case 4: return launch_ext<R, 4>(A, I, psi, dpsi, x_out, s);
    case 5: return launch_ext<R, 5>(A, I, psi, dpsi, x_out, s);
    case 6: return launch_ext<R, 6>(A, I, psi, dpsi, x_out, s);
    case 7: return launch_ext<R, 7>(A, I, psi, dpsi, x_out, s);
    case 8: return launch_ext<R, 8>(A, I, psi, dpsi, x_out, s);
    default: return kUnsupported;
  }
}

template int ext_tables<PNFFT_B_REAL>(const BArgs<PNFFT_B_REAL>&, const IntpolTabs<PNFFT_B_REAL>&, PNFFT_B_REAL*,
                                      PNFFT_B_REAL*, PNFFT_B_REAL*, hipStream_t);

}  // namespace pnfft_amd
