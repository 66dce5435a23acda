// fft.hip -- pruned, shifted, unnormalized 3-D c2c FFT of the NFFT pipeline
// (the PFFT plans of kernel/ndft-parallel.c:964-986, executed at :1524-1546).
//
// Forward (trafo): g[l] = sum_{k in [-N/2,N/2)^3} ghat[k] e^{-2 pi i k.l / n},
// evaluated for every l mod n (natural DFT order on the device grid).
// Backward (adj): ghat[k] = sum_l g[l] e^{+2 pi i k.l / n} for k in [-N/2,N/2)^3.
//
// One HIP kernel per axis pass.  A block stages C pencils of length L in LDS
// (layout [position][pencil], pitch C + 1), runs a Stockham radix-8/4/2 FFT in
// place with one barrier pair per stage, and writes the pencils back.  The
// zero padding N -> n and the k -> k mod n shift are folded into the loads
// (forward) or the stores (backward), so the padded input never exists in HBM.
// Matrix D is fused into the first forward pass and D^T into the last backward
// pass (kernel/matrix_D.c:227-280), with the per-axis factors kept in FP64.
//
// Pass order, forward: axis 2 (contiguous rows) N0*N1 rows  -> buf1[N0][N1][n2]
//                      axis 1 (columns)  N0 x n2 pencils     -> buf2[N0][n1][n2]
//                      axis 0 (columns)  n1 x n2 pencils     -> grid[n0][n1][n2]
// backward runs the mirror image.  Every pass reads and writes each element
// once: the HBM traffic of a pass is (elements in + elements out) * sizeof(C).
#include "device_common.hpp"
#include "fft.hpp"

#include <cmath>

namespace pnfft_amd {

namespace {

constexpr int kThreads = 256;

template <class R> struct FftEpt;
template <> struct FftEpt<float> { static constexpr int value = 32; };   // C*L = 8192
template <> struct FftEpt<double> { static constexpr int value = 16; };  // C*L = 4096

template <class R>
struct PassArgs {
  const Cplx<R>* in;
  Cplx<R>* out;
  long in_outer, in_elem;    // strides in complex elements
  long out_outer, out_elem;
  int n_inner;               // pencils along the contiguous batch index
  int L, Lin, Lout;
  int C;                     // pencils per block
  int nst;
  int radix[16];
  const Cplx<R>* tw;         // tw[i] = e^{-2 pi i i / L}
  int backward;
  // fused deconvolution: factor tab0[k0]*tab1[k1]*tab2[k2]; row r = i0*N1 + i1
  const double* tab0;
  const double* tab1;
  const double* tab2;
  int N1;
  int accumulate;            // backward store: out += value (D^T onto f_hat)
  // frequency <-> index map of the pruned side: index e holds frequency k = e - h;
  // forward loads take k in [-h, pos); the c2r half spectrum is h = N2/2, pos = 1
  // (k2 in [-N2/2, 0]) and its k2 < 0 entries enter the pencil twice (weight 2)
  int h, pos;
  int half;
};

template <class R>
__device__ inline void dft2(Cplx<R>& a, Cplx<R>& b) {
  Cplx<R> t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

// multiply by -i
template <class R>
__device__ inline Cplx<R> mul_mi(Cplx<R> a) { return cmk<R>(a.y, -a.x); }

template <class R>
__device__ inline void dft4(Cplx<R>& v0, Cplx<R>& v1, Cplx<R>& v2, Cplx<R>& v3) {
  Cplx<R> t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_mi(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <class R>
__device__ inline void dft8(Cplx<R>* v) {
  Cplx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  Cplx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  const R h = R(0.70710678118654752440);
  o1 = cmk<R>(h * (o1.x + o1.y), h * (o1.y - o1.x));   // * e^{-i pi/4}
  o2 = mul_mi(o2);                                      // * -i
  o3 = cmk<R>(h * (o3.y - o3.x), -h * (o3.x + o3.y));  // * e^{-3 i pi/4}
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

// One Stockham stage of radix RR over the C pencils held in LDS.  Every thread
// owns exactly EPT/RR butterflies (C*L == EPT*blockDim by construction).
template <class R, int RR, int EPT>
__device__ inline void stage(Cplx<R>* sh, int LD, int C, int L, int Ns, const Cplx<R>* __restrict__ tw) {
  constexpr int NB = EPT / RR;
  const int nbf = L / RR;
  const int tid = threadIdx.x;
  Cplx<R> v[EPT];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int bf = tid + i * kThreads;
    const int c = bf % C, j = bf / C;
#pragma unroll
    for (int q = 0; q < RR; q++) v[i * RR + q] = sh[(j + q * nbf) * LD + c];
  }
  __syncthreads();
  const int twstep = L / (Ns * RR);
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int bf = tid + i * kThreads;
    const int c = bf % C, j = bf / C;
    const int k = j % Ns;
    Cplx<R>* w = v + i * RR;
    if (Ns > 1) {
#pragma unroll
      for (int q = 1; q < RR; q++) w[q] = cmul(w[q], tw[k * q * twstep]);
    }
    if (RR == 2) dft2(w[0], w[1]);
    if (RR == 4) dft4(w[0], w[1], w[2], w[3]);
    if (RR == 8) dft8(w);
    const int base = (j - k) * RR + k;
#pragma unroll
    for (int q = 0; q < RR; q++) sh[(base + q * Ns) * LD + c] = w[q];
  }
  __syncthreads();
}

template <class R>
__device__ inline void run_stages(Cplx<R>* sh, int LD, const PassArgs<R>& P) {
  constexpr int EPT = FftEpt<R>::value;
  int Ns = 1;
  for (int st = 0; st < P.nst; st++) {
    const int rr = P.radix[st];
    if (rr == 8) stage<R, 8, EPT>(sh, LD, P.C, P.L, Ns, P.tw);
    else if (rr == 4) stage<R, 4, EPT>(sh, LD, P.C, P.L, Ns, P.tw);
    else stage<R, 2, EPT>(sh, LD, P.C, P.L, Ns, P.tw);
    Ns *= rr;
  }
}

// position p of a length-L pencil holds frequency k = p (p small) or p - L;
// input index e = k + h when k lies in [-h, pos).
__device__ inline int pruned_src(int p, int L, int h, int pos) {
  if (p < pos) return p + h;
  if (p >= L - h) return p - L + h;
  return -1;
}

// ROWS: pencils are contiguous rows (axis 2); else pencils are strided columns
// whose batch index is contiguous (axes 0 and 1).
template <class R, bool ROWS>
__global__ __launch_bounds__(kThreads) void fft_pass(PassArgs<R> P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* sh = reinterpret_cast<Cplx<R>*>(smem_raw);
  const int C = P.C, L = P.L, LD = C + 1;
  const int tid = threadIdx.x;
  const long outer = blockIdx.y;
  const int c0 = blockIdx.x * C;
  const bool fwd = !P.backward;
  const bool deconv = (P.tab0 != nullptr);

  // Every thread owns exactly EPT = L*C/blockDim positions of the block's
  // pencils; all of its global loads are issued before the LDS writes (and all
  // LDS reads before the global stores) to keep many requests in flight.
  constexpr int EPT = FftEpt<R>::value;
  Cplx<R> v[EPT];

  // ---- load
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int idx = tid + i * kThreads;
    int p, c;
    if (ROWS) { c = idx / L; p = idx % L; }
    else { p = idx / C; c = idx % C; }
    const int col = c0 + c;
    v[i] = cmk<R>(R(0), R(0));
    const int e = fwd ? pruned_src(p, L, P.h, P.pos) : p;
    if (col < P.n_inner && e >= 0) {
      long addr;
      if (ROWS) addr = (long)col * P.in_outer + (long)e * P.in_elem;
      else addr = outer * P.in_outer + col + (long)e * P.in_elem;
      v[i] = P.in[addr];
    }
  }
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int idx = tid + i * kThreads;
    int p, c;
    if (ROWS) { c = idx / L; p = idx % L; }
    else { p = idx / C; c = idx % C; }
    Cplx<R> w = v[i];
    const int col = c0 + c;
    const int e = fwd ? pruned_src(p, L, P.h, P.pos) : p;
    if (deconv && fwd && col < P.n_inner && e >= 0) {
      // row = i0*N1 + i1 (ROWS pass over f_hat), e = i2
      const int i0 = col / P.N1, i1 = col % P.N1;
      const double f = P.tab0[i0] * P.tab1[i1] * P.tab2[e];
      w = cmk<R>(R((double)w.x * f), R((double)w.y * f));
    }
    if (P.half && fwd && e >= 0 && e < P.h) w = cmk<R>(R(2) * w.x, R(2) * w.y);
    if (!fwd) w = cconj(w);
    sh[p * LD + c] = w;
  }
  __syncthreads();

  run_stages<R>(sh, LD, P);

  // ---- store
  const int nstore = P.Lout * C;
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int idx = tid + i * kThreads;
    if (idx < nstore) {
      int e, c;
      if (ROWS) { c = idx / P.Lout; e = idx % P.Lout; }
      else { e = idx / C; c = idx % C; }
      int p = e;
      if (!fwd) p = (e - P.h + L) % L;
      v[i] = sh[p * LD + c];
    }
  }
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int idx = tid + i * kThreads;
    if (idx >= nstore) continue;
    int e, c;
    if (ROWS) { c = idx / P.Lout; e = idx % P.Lout; }
    else { e = idx / C; c = idx % C; }
    const int col = c0 + c;
    if (col >= P.n_inner) continue;
    Cplx<R> w = v[i];
    if (!fwd) w = cconj(w);
    long addr;
    if (ROWS) addr = (long)col * P.out_outer + (long)e * P.out_elem;
    else addr = outer * P.out_outer + col + (long)e * P.out_elem;
    if (deconv && !fwd) {
      const int i0 = col / P.N1, i1 = col % P.N1;
      const double f = P.tab0[i0] * P.tab1[i1] * P.tab2[e];
      w = cmk<R>(R((double)w.x * f), R((double)w.y * f));
    }
    if (P.accumulate) w = cadd(P.out[addr], w);
    P.out[addr] = w;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool fft_length_supported(int L, bool is_double) {
  if (L < 2 || (L & (L - 1))) return false;
  const int cap = is_double ? FftEpt<double>::value * kThreads : FftEpt<float>::value * kThreads;
  return L <= cap;
}

template <class R>
int FftAxis<R>::init(int L_) {
  L = L_;
  if (!fft_length_supported(L, sizeof(R) == 8)) return kUnsupported;
  // radices: as many 8s as possible, then one 4 or 2
  int e = 0;
  while ((1 << e) < L) e++;
  nst = 0;
  int r3 = e % 3;
  for (int i = 0; i < e / 3; i++) radix[nst++] = 8;
  if (r3 == 2) radix[nst++] = 4;
  if (r3 == 1) radix[nst++] = 2;
  C = FftEpt<R>::value * kThreads / L;
  std::vector<Cplx<R>> h(L);
  for (int i = 0; i < L; i++) {
    // angle -2 pi i / L evaluated in FP64
    const double a = -2.0 * 3.14159265358979323846 * (double)i / (double)L;
    h[i] = cmk<R>(R(cos(a)), R(sin(a)));
  }
  if (hipMalloc(&tw, sizeof(Cplx<R>) * L) != hipSuccess) return kHipError;
  if (hipMemcpy(tw, h.data(), sizeof(Cplx<R>) * L, hipMemcpyHostToDevice) != hipSuccess) return kHipError;
  return kOk;
}

template <class R>
void FftAxis<R>::release() {
  if (tw) (void)hipFree(tw);
  tw = nullptr;
}

template <class R>
static PassArgs<R> make_args(const FftAxis<R>& ax) {
  PassArgs<R> P{};
  P.L = ax.L;
  P.C = ax.C;
  P.nst = ax.nst;
  for (int i = 0; i < ax.nst; i++) P.radix[i] = ax.radix[i];
  P.tw = ax.tw;
  return P;
}

template <class R>
static size_t lds_bytes(const FftAxis<R>& ax) {
  return sizeof(Cplx<R>) * (size_t)ax.L * (size_t)(ax.C + 1);
}

template <class R>
int fft3d_forward(const Fft3d<R>& F, const Cplx<R>* fhat, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                  Cplx<R>* grid, hipStream_t s, bool half) {
  const int N0 = F.N[0], N1 = F.N[1], N2 = F.N[2];
  const int n0 = F.n[0], n1 = F.n[1], n2 = F.n[2];
  const int N2in = half ? N2 / 2 + 1 : N2;  // stored length of f_hat rows
  // axis 2: rows (i0,i1) of f_hat -> buf1[i0][i1][p2]
  {
    PassArgs<R> P = make_args(F.ax[2]);
    P.in = fhat; P.out = buf1;
    P.in_outer = N2in; P.in_elem = 1;
    P.out_outer = n2; P.out_elem = 1;
    P.n_inner = N0 * N1;
    P.Lin = N2in; P.Lout = n2;
    P.h = N2 / 2;
    P.pos = half ? 1 : N2 - N2 / 2;
    P.half = half ? 1 : 0;
    if (tab) { P.tab0 = tab; P.tab1 = tab + N0; P.tab2 = tab + N0 + N1; }
    P.N1 = N1;
    dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
    hipLaunchKernelGGL((fft_pass<R, true>), grid_dim, dim3(kThreads), lds_bytes(F.ax[2]), s, P);
  }
  // axis 1: pencils (i0, p2) stride n2 -> buf2[i0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[1]);
    P.in = buf1; P.out = buf2;
    P.in_outer = (long)N1 * n2; P.in_elem = n2;
    P.out_outer = (long)n1 * n2; P.out_elem = n2;
    P.n_inner = n2;
    P.Lin = N1; P.Lout = n1;
    P.h = N1 / 2; P.pos = N1 - N1 / 2;
    dim3 grid_dim((n2 + P.C - 1) / P.C, N0);
    hipLaunchKernelGGL((fft_pass<R, false>), grid_dim, dim3(kThreads), lds_bytes(F.ax[1]), s, P);
  }
  // axis 0: pencils (p1, p2) stride n1*n2 -> grid[p0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[0]);
    P.in = buf2; P.out = grid;
    P.in_outer = n2; P.in_elem = (long)n1 * n2;
    P.out_outer = n2; P.out_elem = (long)n1 * n2;
    P.n_inner = n2;
    P.Lin = N0; P.Lout = n0;
    P.h = N0 / 2; P.pos = N0 - N0 / 2;
    dim3 grid_dim((n2 + P.C - 1) / P.C, n1);
    hipLaunchKernelGGL((fft_pass<R, false>), grid_dim, dim3(kThreads), lds_bytes(F.ax[0]), s, P);
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int fft3d_backward(const Fft3d<R>& F, const Cplx<R>* grid, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                   Cplx<R>* fhat, int accumulate, hipStream_t s, bool half) {
  const int N0 = F.N[0], N1 = F.N[1], N2 = F.N[2];
  const int n0 = F.n[0], n1 = F.n[1], n2 = F.n[2];
  const int N2out = half ? N2 / 2 + 1 : N2;
  // axis 0: grid pencils (p1,p2) -> buf2[i0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[0]);
    P.in = grid; P.out = buf2;
    P.in_outer = n2; P.in_elem = (long)n1 * n2;
    P.out_outer = n2; P.out_elem = (long)n1 * n2;
    P.n_inner = n2;
    P.Lin = n0; P.Lout = N0; P.h = N0 / 2;
    P.backward = 1;
    dim3 grid_dim((n2 + P.C - 1) / P.C, n1);
    hipLaunchKernelGGL((fft_pass<R, false>), grid_dim, dim3(kThreads), lds_bytes(F.ax[0]), s, P);
  }
  // axis 1: buf2[i0][p1][p2] -> buf1[i0][i1][p2]
  {
    PassArgs<R> P = make_args(F.ax[1]);
    P.in = buf2; P.out = buf1;
    P.in_outer = (long)n1 * n2; P.in_elem = n2;
    P.out_outer = (long)N1 * n2; P.out_elem = n2;
    P.n_inner = n2;
    P.Lin = n1; P.Lout = N1; P.h = N1 / 2;
    P.backward = 1;
    dim3 grid_dim((n2 + P.C - 1) / P.C, N0);
    hipLaunchKernelGGL((fft_pass<R, false>), grid_dim, dim3(kThreads), lds_bytes(F.ax[1]), s, P);
  }
  // axis 2: rows buf1[i0][i1][:] -> f_hat[i0][i1][:] (D^T fused; c2r keeps k2 in [-N2/2, 0])
  {
    PassArgs<R> P = make_args(F.ax[2]);
    P.in = buf1; P.out = fhat;
    P.in_outer = n2; P.in_elem = 1;
    P.out_outer = N2out; P.out_elem = 1;
    P.n_inner = N0 * N1;
    P.Lin = n2; P.Lout = N2out; P.h = N2 / 2;
    P.backward = 1;
    P.accumulate = accumulate;
    if (tab) { P.tab0 = tab; P.tab1 = tab + N0; P.tab2 = tab + N0 + N1; }
    P.N1 = N1;
    dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
    hipLaunchKernelGGL((fft_pass<R, true>), grid_dim, dim3(kThreads), lds_bytes(F.ax[2]), s, P);
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int Fft3d<R>::init(const int* N_, const int* n_) {
  for (int t = 0; t < 3; t++) {
    N[t] = N_[t];
    n[t] = n_[t];
    if (N[t] > n[t]) return kUnsupported;
    int st = ax[t].init(n[t]);
    if (st != kOk) return st;
  }
  for (int t = 0; t < 3; t++) {
    const size_t bytes = sizeof(Cplx<R>) * (size_t)ax[t].L * (size_t)(ax[t].C + 1);
    if (bytes > 64 * 1024) {
      if (hipFuncSetAttribute((const void*)fft_pass<R, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes) != hipSuccess)
        return kHipError;
      if (hipFuncSetAttribute((const void*)fft_pass<R, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes) != hipSuccess)
        return kHipError;
    }
  }
  return kOk;
}

template <class R>
void Fft3d<R>::release() {
  for (int t = 0; t < 3; t++) ax[t].release();
}

template struct FftAxis<float>;
template struct FftAxis<double>;
template struct Fft3d<float>;
template struct Fft3d<double>;
template int fft3d_forward<float>(const Fft3d<float>&, const Cplx<float>*, const double*, Cplx<float>*,
                                  Cplx<float>*, Cplx<float>*, hipStream_t, bool);
template int fft3d_forward<double>(const Fft3d<double>&, const Cplx<double>*, const double*, Cplx<double>*,
                                   Cplx<double>*, Cplx<double>*, hipStream_t, bool);
template int fft3d_backward<float>(const Fft3d<float>&, const Cplx<float>*, const double*, Cplx<float>*,
                                   Cplx<float>*, Cplx<float>*, int, hipStream_t, bool);
template int fft3d_backward<double>(const Fft3d<double>&, const Cplx<double>*, const double*, Cplx<double>*,
                                    Cplx<double>*, Cplx<double>*, int, hipStream_t, bool);

}  // namespace pnfft_amd
