// fft.hip -- pruned, shifted, unnormalized 3-D c2c FFT of the NFFT pipeline
// (the PFFT plans of kernel/ndft-parallel.c:964-986, executed at :1524-1546).
//
// Forward (trafo): g[l] = sum_{k in [-N/2,N/2)^3} ghat[k] e^{-2 pi i k.l / n},
// evaluated for every l mod n (natural DFT order on the device grid).
// Backward (adj): ghat[k] = sum_l g[l] e^{+2 pi i k.l / n} for k in [-N/2,N/2)^3.
//
// One HIP kernel per axis pass, instantiated per pencil length L so that every
// index is a compile-time shift or mask.  A block holds C pencils of length L
// (layout [position][pencil], pitch C + 1) and runs a Stockham radix-8/4/2 FFT
// with one barrier pair per LDS stage.  The column passes compute their first
// stage on the loaded registers and their last stage straight into the stores,
// so a 512-point pencil makes one LDS round trip instead of four.  The
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
#include <type_traits>

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
  const Cplx<R>* tw;         // tw[i] = e^{-2 pi i i / L}
  int backward;
  // fused deconvolution: factor tab0[k0]*tab1[k1]*tab2[k2]; row r = i0*N1 + i1
  const double* tab0;
  const double* tab1;
  const double* tab2;
  int N1;
  int ctab;                  // tab0..2 hold complex factors (re, im pairs) per frequency index
  int accumulate;            // backward store: out += value (D^T onto f_hat)
  int colfac;                // columns pass over f_hat's axis 1 (c2r plans): factors at (plane, position, column)
  // frequency <-> index map of the pruned side: index e holds frequency k = e - h;
  // forward loads take k in [-h, pos)
  int h, pos;
  // c2r / r2c row passes: twiddles tw2[p] = e^{-i pi p / L} of the length-2L real
  // transform and N2 / 2 (the stored half spectrum k2 in [-N2/2, 0], N2/2 + 1 values)
  const Cplx<R>* tw2;
  int N2h;
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

// Block shape of a pass over pencils of length L with NT threads: C pencils per
// block with C * L = EPT * NT, LDS layout [position][pencil] with pitch LD.
template <class R, int L, int NT = kThreads, int EPT_ = FftEpt<R>::value>
struct PassShape {
  static constexpr int EPT = EPT_;
  static constexpr int C = EPT * NT / L;
  static constexpr int LD = C == 1 ? 1 : C + 1;
  static constexpr size_t LDS = sizeof(Cplx<R>) * (size_t)L * (size_t)LD;
};

// Threads of a columns pass: 512 where the doubled tile still fits in 160 KB of
// LDS (twice the pencils per block, so every row segment is twice as long:
// 256 B at L = 512 in FP32), else 256.
template <class R, int L>
constexpr int cols_threads() {
  return PassShape<R, L, 2 * kThreads>::LDS <= 160 * 1024 ? 2 * kThreads : kThreads;
}

// Elements per thread of a columns pass.  The narrow variant, for the axis-1
// pass (pencil elements n2 apart), takes 16 for 512-point FP32 pencils: 16
// pencils per 512-thread block, 64 KB of LDS, two blocks per CU.  The axis-0
// pass (elements n1 n2 apart) keeps 32: its 256-byte row segments matter more
// there than the second block.
template <class R, int L, bool NARROW>
constexpr int cols_ept() {
  return (NARROW && sizeof(R) == 4 && L == 512) ? 16 : FftEpt<R>::value;
}

template <class R, int L, bool NARROW>
using ColsShape = PassShape<R, L, cols_threads<R, L>(), cols_ept<R, L, NARROW>()>;

// The wide 512-point FP32 pass (32 pencils, 512 threads) makes its two LDS
// round trips in two halves of the pencil positions, so a block holds 256
// positions (68 KB) and a CU two blocks: one block's loads and stores overlap
// the other's LDS work.
template <class R, int L, bool NARROW, bool FAC = false>
constexpr bool cols_split() {
  using S = ColsShape<R, L, NARROW>;
  return !FAC && !NARROW && sizeof(R) == 4 && L == 512 && S::C == 32 && cols_threads<R, L>() == 512 &&
         S::EPT == 32;
}

template <class R, int L, bool NARROW, bool FAC = false>
constexpr size_t cols_lds() {
  using S = ColsShape<R, L, NARROW>;
  return cols_split<R, L, NARROW, FAC>() ? S::LDS / 2 : S::LDS;
}

// radix of the stage that starts at span Ns (8 while possible, then 4 or 2)
template <int L, int Ns>
struct StageRadix {
  static constexpr int value = (L / Ns) >= 8 ? 8 : (L / Ns);
};

template <class R, int RR>
__device__ inline void dft(Cplx<R>* w) {
  if constexpr (RR == 2) dft2(w[0], w[1]);
  if constexpr (RR == 4) dft4(w[0], w[1], w[2], w[3]);
  if constexpr (RR == 8) dft8(w);
}

// One Stockham stage LDS -> LDS of radix RR at span Ns.  Butterfly bf of the
// block covers pencil c = bf % C, index j = bf / C; it reads positions
// j + q L/RR and writes (j - k) RR + k + q Ns with k = j mod Ns.
template <class R, int L, int Ns, int NT, int EP>
__device__ inline void stage_lds(Cplx<R>* sh, const Cplx<R>* __restrict__ tw) {
  using S = PassShape<R, L, NT, EP>;
  constexpr int RR = StageRadix<L, Ns>::value;
  constexpr int NB = S::EPT / RR;
  constexpr int nbf = L / RR;
  constexpr int twstep = L / (Ns * RR);
  const int tid = threadIdx.x;
  Cplx<R> v[S::EPT];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int bf = tid + i * NT;
    const int c = bf % S::C, j = bf / S::C;
#pragma unroll
    for (int q = 0; q < RR; q++) v[i * RR + q] = sh[(j + q * nbf) * S::LD + c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int bf = tid + i * NT;
    const int c = bf % S::C, j = bf / S::C;
    const int k = j % Ns;
    Cplx<R>* w = v + i * RR;
    if constexpr (Ns > 1) {
#pragma unroll
      for (int q = 1; q < RR; q++) w[q] = cmul(w[q], tw[k * q * twstep]);
    }
    dft<R, RR>(w);
    const int base = (j - k) * RR + k;
#pragma unroll
    for (int q = 0; q < RR; q++) sh[(base + q * Ns) * S::LD + c] = w[q];
  }
  __syncthreads();
}

// stages LDS -> LDS from span Ns up to, but not including, span END
template <class R, int L, int Ns, int END, int NT, int EP>
__device__ inline void stages_lds(Cplx<R>* sh, const Cplx<R>* __restrict__ tw) {
  if constexpr (Ns < END) {
    stage_lds<R, L, Ns, NT, EP>(sh, tw);
    stages_lds<R, L, Ns * StageRadix<L, Ns>::value, END, NT, EP>(sh, tw);
  }
}

// span of the last stage: Ns with Ns * radix(Ns) == L
template <int L, int Ns = 1>
constexpr int last_span() {
  if constexpr (Ns * StageRadix<L, Ns>::value >= L) return Ns;
  else return last_span<L, Ns * StageRadix<L, Ns>::value>();
}

// ---- per-element transforms of the loads and stores ----------------------

// position p of a length-L pencil holds frequency k = p (p small) or p - L;
// input index e = k + h when k lies in [-h, pos).
__device__ inline int pruned_src(int p, int L, int h, int pos) {
  if (p < pos) return p + h;
  if (p >= L - h) return p - L + h;
  return -1;
}

// forward load of position p: the pruned input index e (or -1 for a zero)
template <class R>
__device__ inline int load_index(const PassArgs<R>& P, int p, int L) {
  return P.backward ? p : pruned_src(p, L, P.h, P.pos);
}

// w times the product of the three per-axis factors of frequency (i0, i1, e),
// formed and applied in FP64: real factors, or complex ones (P.ctab) for the
// phases and derivative factors of interlaced and ik-differentiated plans
template <class R>
__device__ inline Cplx<R> apply_factors(const PassArgs<R>& P, Cplx<R> w, int i0, int i1, int e) {
  if (!P.ctab) {
    const double f = P.tab0[i0] * P.tab1[i1] * P.tab2[e];
    return cmk<R>(R((double)w.x * f), R((double)w.y * f));
  }
  const double ar = P.tab0[2 * i0], ai = P.tab0[2 * i0 + 1];
  const double br = P.tab1[2 * i1], bi = P.tab1[2 * i1 + 1];
  const double cr = P.tab2[2 * e], ci = P.tab2[2 * e + 1];
  const double abr = ar * br - ai * bi, abi = ar * bi + ai * br;
  const double fr = abr * cr - abi * ci, fi = abr * ci + abi * cr;
  const double wr = (double)w.x, wi = (double)w.y;
  return cmk<R>(R(wr * fr - wi * fi), R(wr * fi + wi * fr));
}

// value as it enters the transform: D factor (forward rows), conjugation (backward)
template <class R>
__device__ inline Cplx<R> on_load(const PassArgs<R>& P, Cplx<R> w, int col, int e) {
  const bool fwd = !P.backward;
  if (fwd && P.tab0 != nullptr) {
    // row = i0*N1 + i1 (rows pass over f_hat), e = i2
    w = apply_factors(P, w, col / P.N1, col % P.N1, e);
  }
  if (!fwd) w = cconj(w);
  return w;
}

// value as it leaves: conjugation and D^T factor (backward), accumulation
template <class R>
__device__ inline void on_store(const PassArgs<R>& P, Cplx<R> w, int col, int e, long addr) {
  if (P.backward) {
    w = cconj(w);
    if (P.tab0 != nullptr) w = apply_factors(P, w, col / P.N1, col % P.N1, e);
  }
  if (P.accumulate) w = cadd(P.out[addr], w);
  P.out[addr] = w;
}

// output position p of a pencil -> stored index e, or -1 (pruned away)
template <class R>
__device__ inline int store_index(const PassArgs<R>& P, int p, int L) {
  if (!P.backward) return p;
  int e = p + P.h;
  if (e >= L) e -= L;
  return e < P.Lout ? e : -1;
}

// Rows pass (axis 2, contiguous pencils): loads and stores walk each row with
// consecutive lanes; all stages run in LDS.
template <class R, int L>
__global__ __launch_bounds__(kThreads) void fft_rows(PassArgs<R> P) {
  using S = PassShape<R, L>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* sh = reinterpret_cast<Cplx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * S::C;
  Cplx<R> v[S::EPT];
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int col = c0 + c;
    const int e = load_index(P, p, L);
    v[i] = cmk<R>(R(0), R(0));
    if (col < P.n_inner && e >= 0) v[i] = P.in[(long)col * P.in_outer + e];
  }
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int col = c0 + c;
    const int e = load_index(P, p, L);
    Cplx<R> w = v[i];
    if (col < P.n_inner && e >= 0) w = on_load(P, w, col, e);
    sh[p * S::LD + c] = w;
  }
  __syncthreads();
  stages_lds<R, L, 1, L, kThreads, S::EPT>(sh, P.tw);
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    v[i] = sh[p * S::LD + c];
  }
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int col = c0 + c;
    const int e = store_index(P, p, L);
    if (col < P.n_inner && e >= 0) on_store(P, v[i], col, e, (long)col * P.out_outer + e);
  }
}

// Coefficient S[k] (k = 0..L) of the Hermitian spectrum of a real row of length
// 2L whose stored half u holds the frequencies k2 = e - Nh in [-Nh, 0]: the row is
// Re(u[Nh] + 2 sum_{k2 < 0} u[k2 + Nh] e^{-2 pi i k2 l / 2L}), i.e. S[0] = Re u[Nh],
// S[k] = u[Nh - k] for 0 < k < L (zero beyond Nh), and at k = L (only when
// Nh = L) the Nyquist value 2 Re u[0].
template <class R>
__device__ inline Cplx<R> half_coef(const Cplx<R>* u, int k, int Nh, int L) {
  if (k == 0) return cmk<R>(u[Nh].x, R(0));
  if (k == L) return Nh == L ? cmk<R>(R(2) * u[0].x, R(0)) : cmk<R>(R(0), R(0));
  return k <= Nh ? u[Nh - k] : cmk<R>(R(0), R(0));
}

// c2r rows pass (axis 2 of a c2r plan, last forward pass): the real row x of
// length 2L from the half spectrum u (Nh + 1 values per row) as one complex
// transform of length L: z = IDFT_L(E + i O) with E[k] = S[k] + conj S[L - k] and
// O[k] = (S[k] - conj S[L - k]) e^{i pi k / L} gives z[j] = x[2j] + i x[2j + 1],
// so the complex output row is the real row in memory.  The inverse transform
// is conj(FFT(conj)).
template <class R, int L>
__global__ __launch_bounds__(kThreads) void fft_rows_c2r(PassArgs<R> P) {
  using S = PassShape<R, L>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* sh = reinterpret_cast<Cplx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * S::C;
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int row = c0 + c;
    Cplx<R> z = cmk<R>(R(0), R(0));
    if (row < P.n_inner) {
      const Cplx<R>* u = P.in + (long)row * P.in_outer;
      const Cplx<R> a = half_coef<R>(u, p, P.N2h, L), b = cconj(half_coef<R>(u, L - p, P.N2h, L));
      const Cplx<R> E = cadd(a, b), O = cmul(csub(a, b), cconj(P.tw2[p]));
      z = cmk<R>(E.x - O.y, E.y + O.x);
    }
    sh[p * S::LD + c] = cconj(z);
  }
  __syncthreads();
  stages_lds<R, L, 1, L, kThreads, S::EPT>(sh, P.tw);
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int row = c0 + c;
    if (row < P.n_inner) P.out[(long)row * P.out_outer + p] = cconj(sh[p * S::LD + c]);
  }
}

// r2c rows pass (axis 2 of a c2r plan, first backward pass): for the real row x
// of length 2L read as z[j] = x[2j] + i x[2j + 1], Z = DFT+_L(z) (the adjoint's
// sign, conj(FFT(conj))) gives G[k] = sum_l x[l] e^{+2 pi i k l / 2L} =
// (Z[k] + conj Z[L - k]) / 2 + e^{i pi k / L} (Z[k] - conj Z[L - k]) / 2i, and the
// stored half k2 = -k in [-Nh, 0] is conj G[k].
template <class R, int L>
__global__ __launch_bounds__(kThreads) void fft_rows_r2c(PassArgs<R> P) {
  using S = PassShape<R, L>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* sh = reinterpret_cast<Cplx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * S::C;
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int row = c0 + c;
    Cplx<R> z = cmk<R>(R(0), R(0));
    if (row < P.n_inner) z = P.in[(long)row * P.in_outer + p];
    sh[p * S::LD + c] = cconj(z);
  }
  __syncthreads();
  stages_lds<R, L, 1, L, kThreads, S::EPT>(sh, P.tw);
  const int Nh = P.N2h;
#pragma unroll
  for (int i = 0; i < S::EPT; i++) {
    const int idx = tid + i * kThreads;
    const int c = idx / L, p = idx % L;
    const int row = c0 + c;
    if (row < P.n_inner && p <= Nh) {
      const Cplx<R> zk = cconj(sh[p * S::LD + c]);
      const Cplx<R> zm = sh[((L - p) & (L - 1)) * S::LD + c];  // conj Z[L - k]
      const Cplx<R> ge = cmk<R>(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y + zm.y));
      const Cplx<R> d = csub(zk, zm);
      const Cplx<R> go = cmk<R>(R(0.5) * d.y, R(-0.5) * d.x);  // d / 2i
      const Cplx<R> G = cadd(ge, cmul(cconj(P.tw2[p]), go));
      Cplx<R>* o = P.out + (long)row * P.out_outer;
      o[Nh - p] = cconj(G);
      if (p == 0 && Nh == L) o[0] = cmk<R>(zk.x - zk.y, R(0));  // k = L: G = Re Z[0] - Im Z[0]
    }
  }
}

// LDS writes of one lane become visible to the other lanes of its wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows pass for L = 512 = 8^3: one wave per pencil (four per block).  Lane j
// holds butterfly j of every stage: the first stage runs on the loaded
// registers (positions j + 64 q, consecutive lanes on consecutive elements),
// the middle stage goes through the wave's LDS pencil (element p at p + p / 8,
// which keeps the stage's reads and writes within two-way bank conflicts), and
// the last stage's outputs j + 64 q go straight to HBM.  Same Stockham stages,
// twiddles and per-element hooks as fft_rows.  Everything a wave reads from
// global memory is requested before its first stage: in FP32 the block stages
// the twiddles in LDS (one barrier), and a backward pass with real factors
// loads its D^T factors and, when it accumulates, the f_hat values it adds to
// up front, so a pencil costs one memory round trip instead of three or four.
constexpr int kRows512Pencils = 4;

// PRE: a backward pass with real factors or accumulation (the launch checks it)
template <class R, bool PRE>
__global__ __launch_bounds__(256) void fft_rows512(PassArgs<R> P) {
  constexpr int L = 512, NP = L + L / 8;
  constexpr bool TW_LDS = sizeof(R) == 4;
  constexpr int NTW = 448;  // twiddles used: tw[8 k q] (k, q < 8) and tw[j q] (j < 64, q < 8), all below 448
  __shared__ Cplx<R> sh[kRows512Pencils][NP];
  __shared__ Cplx<R> stw[TW_LDS ? NTW : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = blockIdx.x * kRows512Pencils + wave;
  const bool live = col < P.n_inner;
  Cplx<R>* pen = sh[wave];
  auto pad = [](int p) { return p + (p >> 3); };
  Cplx<R> v[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = load_index(P, lane + 64 * q, L);
    v[q] = cmk<R>(R(0), R(0));
    if (live && e >= 0) v[q] = P.in[(long)col * P.in_outer + e];
  }
  // backward pass with real factors: D^T factors and the accumulated values, loaded now
  double fac[PRE ? 8 : 1];
  Cplx<R> acc[PRE ? 8 : 1];
  if constexpr (PRE) {
    const double f01 = P.tab0 ? P.tab0[col / P.N1] * P.tab1[col % P.N1] : 1.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = store_index(P, lane + 64 * q, L);
      fac[q] = (P.tab0 && e >= 0) ? f01 * P.tab2[e] : 1.0;
      acc[q] = cmk<R>(R(0), R(0));
      if (live && P.accumulate && e >= 0) acc[q] = P.out[(long)col * P.out_outer + e];
    }
  }
  const Cplx<R>* tw = P.tw;
  if constexpr (TW_LDS) {
    for (int i = threadIdx.x; i < NTW; i += 256) stw[i] = P.tw[i];
    __syncthreads();
    tw = stw;
  }
  if (!live) return;  // whole wave; no barrier follows
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = load_index(P, lane + 64 * q, L);
    if (e >= 0) v[q] = on_load(P, v[q], col, e);
  }
  // stage span 1: outputs at 8 j + q
  dft8(v);
#pragma unroll
  for (int q = 0; q < 8; q++) pen[pad(8 * lane + q)] = v[q];
  wave_sync();
  // stage span 8: inputs j + 64 q, twiddles tw[8 k q] with k = j mod 8, outputs (j - k) 8 + k + 8 q
  {
    const int k = lane & 7;
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = pen[pad(lane + 64 * q)];
#pragma unroll
    for (int q = 1; q < 8; q++) v[q] = cmul(v[q], tw[8 * k * q]);
    dft8(v);
    wave_sync();  // every read of the pencil precedes the writes
    const int base = (lane - k) * 8 + k;
#pragma unroll
    for (int q = 0; q < 8; q++) pen[pad(base + 8 * q)] = v[q];
    wave_sync();
  }
  // stage span 64: inputs j + 64 q, twiddles tw[j q], outputs j + 64 q
#pragma unroll
  for (int q = 0; q < 8; q++) v[q] = pen[pad(lane + 64 * q)];
#pragma unroll
  for (int q = 1; q < 8; q++) v[q] = cmul(v[q], tw[lane * q]);
  dft8(v);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = store_index(P, lane + 64 * q, L);
    if (e < 0) continue;
    const long addr = (long)col * P.out_outer + e;
    if constexpr (PRE) {
      // on_store with the factors and accumulated values loaded above
      Cplx<R> w = cconj(v[q]);
      w = cmk<R>(R((double)w.x * fac[q]), R((double)w.y * fac[q]));
      if (P.accumulate) w = cadd(acc[q], w);
      P.out[addr] = w;
    } else {
      on_store(P, v[q], col, e, addr);
    }
  }
}

// Columns pass (axes 0 and 1): pencil c of the block is column c0 + c of
// plane blockIdx.y, consecutive lanes take consecutive columns.  The first
// stage runs on the loaded registers and the last stage stores straight to
// HBM, so only the middle stages touch LDS.
// FAC: the pass over f_hat's axis 1 of a c2r plan, which takes the per-axis
// factors (D, or the complex ones) of element (plane, position, column) on its
// loads (forward) or its stores (backward, with accumulation onto f_hat)
template <class R, int L, bool NARROW, bool FAC>
__global__ __launch_bounds__((cols_threads<R, L>()), (cols_split<R, L, NARROW, FAC>() ? 4 : 1)) void fft_cols(
    PassArgs<R> P) {
  constexpr int NT = cols_threads<R, L>();
  using S = ColsShape<R, L, NARROW>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* sh = reinterpret_cast<Cplx<R>*>(smem_raw);
  constexpr int R1 = StageRadix<L, 1>::value;
  constexpr int NB1 = S::EPT / R1, nbf1 = L / R1;
  constexpr int NsL = last_span<L>();
  constexpr int RL = StageRadix<L, NsL>::value;
  constexpr int NBL = S::EPT / RL, nbfL = L / RL;
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * S::C;
  // FP32 512-point passes read their twiddles from LDS (staged here; the first
  // block barrier below precedes every use)
  constexpr bool TW_LDS = sizeof(R) == 4 && L == 512;
  __shared__ Cplx<R> stw[TW_LDS ? L : 1];
  const Cplx<R>* tw = P.tw;
  if constexpr (TW_LDS) {
    for (int i = tid; i < L; i += NT) stw[i] = P.tw[i];
    tw = stw;
  }
  const long in_base = (long)blockIdx.y * P.in_outer;
  const long out_base = (long)blockIdx.y * P.out_outer;
  Cplx<R> v[S::EPT];
  // ---- first stage (span 1): butterfly (c, j) takes positions j + q L/R1
  auto load_bf = [&](int i) {
    const int bf = tid + i * NT;
    const int c = bf % S::C, j = bf / S::C;
    const int col = c0 + c;
#pragma unroll
    for (int q = 0; q < R1; q++) {
      const int e = load_index(P, j + q * nbf1, L);
      v[i * R1 + q] = cmk<R>(R(0), R(0));
      if (col < P.n_inner && e >= 0) {
        v[i * R1 + q] = P.in[in_base + col + (long)e * P.in_elem];
        if constexpr (FAC) {
          if (!P.backward && P.tab0) v[i * R1 + q] = apply_factors(P, v[i * R1 + q], (int)blockIdx.y, e, col);
        }
      }
    }
  };
  auto stage1_bf = [&](int i) {
    if (P.backward) {
#pragma unroll
      for (int q = 0; q < R1; q++) v[i * R1 + q] = cconj(v[i * R1 + q]);
    }
    dft<R, R1>(v + i * R1);
  };
  if constexpr (!cols_split<R, L, NARROW, FAC>()) {
#pragma unroll
    for (int i = 0; i < NB1; i++) load_bf(i);
#pragma unroll
    for (int i = 0; i < NB1; i++) stage1_bf(i);
  }
  if constexpr (cols_split<R, L, NARROW, FAC>()) {
    // L = 512 = 8^3, butterflies i = 0..3 of pencil c at j = j0 + 16 i; stage
    // outputs of butterflies 2h, 2h + 1 and stage inputs q = 4h .. 4h + 3 lie in
    // the half h of the positions, [256 h, 256 h + 256).  Butterflies 2 and 3
    // are loaded while half 0 goes through the LDS.
    constexpr int LD = S::LD, HALF = L / 2;
    const int c = tid % S::C, j0 = tid / S::C;
    Cplx<R> u[S::EPT];
    load_bf(0);
    load_bf(1);
    stage1_bf(0);
    stage1_bf(1);
    // stage 1 outputs (8 j + q) -> stage 2 inputs (j + 64 q)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1) {
        stage1_bf(2);
        stage1_bf(3);
      }
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const int i = 2 * h + ii, j = j0 + 16 * i;
#pragma unroll
        for (int q = 0; q < 8; q++) sh[(8 * j + q - HALF * h) * LD + c] = v[i * 8 + q];
      }
      if (h == 0) {
        load_bf(2);
        load_bf(3);
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = j0 + 16 * i;
#pragma unroll
        for (int qq = 0; qq < 4; qq++) {
          const int q = 4 * h + qq;
          u[i * 8 + q] = sh[(j + 64 * q - HALF * h) * LD + c];
        }
      }
      __syncthreads();
    }
    // stage 2 (span 8): twiddles tw[8 k q], k = j mod 8; outputs (j - k) 8 + k + 8 q
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int j = j0 + 16 * i, k = j & 7;
      Cplx<R>* w = u + i * 8;
#pragma unroll
      for (int q = 1; q < 8; q++) w[q] = cmul(w[q], tw[8 * k * q]);
      dft8(w);
    }
    // stage 2 outputs -> stage 3 inputs (j + 64 q)
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const int i = 2 * h + ii, j = j0 + 16 * i, k = j & 7;
#pragma unroll
        for (int q = 0; q < 8; q++) sh[((j - k) * 8 + k + 8 * q - HALF * h) * LD + c] = u[i * 8 + q];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = j0 + 16 * i;
#pragma unroll
        for (int qq = 0; qq < 4; qq++) {
          const int q = 4 * h + qq;
          v[i * 8 + q] = sh[(j + 64 * q - HALF * h) * LD + c];
        }
      }
      __syncthreads();
    }
    // stage 3 (span 64): twiddles tw[j q], outputs j + 64 q stored per butterfly
    // (its twiddles are loaded where they are used, not all at once)
    const int col = c0 + c;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int j = j0 + 16 * i;
      Cplx<R>* w = v + i * 8;
#pragma unroll
      for (int q = 1; q < 8; q++) w[q] = cmul(w[q], tw[j * q]);
      dft8(w);
      if (col < P.n_inner) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int e = store_index(P, j + 64 * q, L);
          if (e >= 0) P.out[out_base + col + (long)e * P.out_elem] = P.backward ? cconj(w[q]) : w[q];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    return;
  } else if constexpr (NsL > 1) {
#pragma unroll
    for (int i = 0; i < NB1; i++) {
      const int bf = tid + i * NT;
      const int c = bf % S::C, j = bf / S::C;
#pragma unroll
      for (int q = 0; q < R1; q++) sh[(j * R1 + q) * S::LD + c] = v[i * R1 + q];
    }
    __syncthreads();
    stages_lds<R, L, R1, NsL, NT, S::EPT>(sh, tw);
    // ---- last stage (span NsL): outputs at positions j + q NsL
#pragma unroll
    for (int i = 0; i < NBL; i++) {
      const int bf = tid + i * NT;
      const int c = bf % S::C, j = bf / S::C;
#pragma unroll
      for (int q = 0; q < RL; q++) v[i * RL + q] = sh[(j + q * nbfL) * S::LD + c];
    }
    constexpr int twstep = L / (NsL * RL);
#pragma unroll
    for (int i = 0; i < NBL; i++) {
      const int bf = tid + i * NT;
      const int j = bf / S::C;
#pragma unroll
      for (int q = 1; q < RL; q++) v[i * RL + q] = cmul(v[i * RL + q], tw[j * q * twstep]);
      dft<R, RL>(v + i * RL);
    }
  }
  // positions of the outputs: j + q NsL (a single stage: NsL = 1, j + q nbf1)
  constexpr int RO = NsL > 1 ? RL : R1;
  constexpr int NBO = S::EPT / RO;
  constexpr int stepO = NsL > 1 ? NsL : nbf1;
#pragma unroll
  for (int i = 0; i < NBO; i++) {
    const int bf = tid + i * NT;
    const int c = bf % S::C, j = bf / S::C;
    const int col = c0 + c;
    if (col >= P.n_inner) continue;
#pragma unroll
    for (int q = 0; q < RO; q++) {
      const int e = store_index(P, j + q * stepO, L);
      if (e >= 0) {
        const long addr = out_base + col + (long)e * P.out_elem;
        Cplx<R> w = P.backward ? cconj(v[i * RO + q]) : v[i * RO + q];
        if constexpr (FAC) {
          // the c2r plan's last backward pass: D^T (or the complex factors) and accumulation onto f_hat
          if (P.backward && P.tab0) w = apply_factors(P, w, (int)blockIdx.y, e, col);
          if (P.accumulate) w = cadd(P.out[addr], w);
        }
        P.out[addr] = w;
      }
    }
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
void FftAxis<R>::release() {
  if (tw) (void)hipFree(tw);
  if (tw2) (void)hipFree(tw2);
  tw = nullptr;
  tw2 = nullptr;
}

template <class R>
static PassArgs<R> make_args(const FftAxis<R>& ax) {
  PassArgs<R> P{};
  P.L = ax.L;
  P.C = ax.C;
  P.tw = ax.tw;
  return P;
}

template <class R, int L>
static int pass_attr() {
  constexpr size_t lr = PassShape<R, L>::LDS;
  constexpr size_t lc = cols_lds<R, L, false>();
  constexpr size_t ln = cols_lds<R, L, true>();
  auto attr = [](const void* f, size_t lds) {
    return lds <= 64 * 1024 ||
           hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
  };
  if (!attr((const void*)fft_rows<R, L>, lr) || !attr((const void*)fft_rows_c2r<R, L>, lr) ||
      !attr((const void*)fft_rows_r2c<R, L>, lr) || !attr((const void*)fft_cols<R, L, false, false>, lc) ||
      !attr((const void*)fft_cols<R, L, true, false>, ln) || !attr((const void*)fft_cols<R, L, true, true>, ln))
    return kHipError;
  return kOk;
}

enum PassKind { kPassRows = 0, kPassCols = 1, kPassRowsC2r = 2, kPassRowsR2c = 3 };

template <class R, int L>
static void pass_launch(int kind, dim3 grid, const PassArgs<R>& P, hipStream_t s) {
  if (kind == kPassRows) {
    if constexpr (L == 512) {
      if (P.backward && (P.tab0 != nullptr || P.accumulate) && !P.ctab)
        hipLaunchKernelGGL((fft_rows512<R, true>), grid, dim3(256), 0, s, P);
      else
        hipLaunchKernelGGL((fft_rows512<R, false>), grid, dim3(256), 0, s, P);
    } else {
      constexpr size_t lds = PassShape<R, L>::LDS;
      hipLaunchKernelGGL((fft_rows<R, L>), grid, dim3(kThreads), lds, s, P);
    }
  } else if (kind == kPassRowsC2r) {
    constexpr size_t lds = PassShape<R, L>::LDS;
    hipLaunchKernelGGL((fft_rows_c2r<R, L>), grid, dim3(kThreads), lds, s, P);
  } else if (kind == kPassRowsR2c) {
    constexpr size_t lds = PassShape<R, L>::LDS;
    hipLaunchKernelGGL((fft_rows_r2c<R, L>), grid, dim3(kThreads), lds, s, P);
  } else {
    constexpr int nt = cols_threads<R, L>();
    constexpr size_t lf = cols_lds<R, L, true, true>();
    constexpr size_t ln = cols_lds<R, L, true>();
    constexpr size_t lw = cols_lds<R, L, false>();
    if (P.colfac) {  // planned with the narrow pencil count
      hipLaunchKernelGGL((fft_cols<R, L, true, true>), grid, dim3(nt), lf, s, P);
    } else if (ColsShape<R, L, true>::C != ColsShape<R, L, false>::C && P.C == ColsShape<R, L, true>::C) {
      // the narrow variant is the one whose pencil count the pass was planned with
      hipLaunchKernelGGL((fft_cols<R, L, true, false>), grid, dim3(nt), ln, s, P);
    } else {
      hipLaunchKernelGGL((fft_cols<R, L, false, false>), grid, dim3(nt), lw, s, P);
    }
  }
}

// Runs F(L) for the supported lengths; F is a generic lambda taking the
// length as std::integral_constant.
template <class R, class F>
static int with_length(int L, F&& f) {
  switch (L) {
#define PNFFT_FFT_L(V) \
  case V: return f(std::integral_constant<int, V>{});
    PNFFT_FFT_L(2)
    PNFFT_FFT_L(4)
    PNFFT_FFT_L(8)
    PNFFT_FFT_L(16)
    PNFFT_FFT_L(32)
    PNFFT_FFT_L(64)
    PNFFT_FFT_L(128)
    PNFFT_FFT_L(256)
    PNFFT_FFT_L(512)
    PNFFT_FFT_L(1024)
    PNFFT_FFT_L(2048)
    PNFFT_FFT_L(4096)
#undef PNFFT_FFT_L
    case 8192:
      if constexpr (sizeof(R) == 4) return f(std::integral_constant<int, 8192>{});
      return kUnsupported;
    default:
      return kUnsupported;
  }
}

template <class R>
int FftAxis<R>::init(int L_, bool half_angles) {
  L = L_;
  if (!fft_length_supported(L, sizeof(R) == 8)) return kUnsupported;
  C = L == 512 ? kRows512Pencils : FftEpt<R>::value * kThreads / L;
  Ccols = 0;
  (void)with_length<R>(L, [&](auto Lc) {
    constexpr int LL = decltype(Lc)::value;
    Ccols = ColsShape<R, LL, false>::C;
    Ccols_narrow = ColsShape<R, LL, true>::C;
    return kOk;
  });
  if (Ccols <= 0) return kUnsupported;
  std::vector<Cplx<R>> h(L);
  for (int i = 0; i < L; i++) {
    // angle -2 pi i / L evaluated in FP64
    const double a = -2.0 * 3.14159265358979323846 * (double)i / (double)L;
    h[i] = cmk<R>(R(cos(a)), R(sin(a)));
  }
  if (hipMalloc(&tw, sizeof(Cplx<R>) * L) != hipSuccess) return kHipError;
  if (hipMemcpy(tw, h.data(), sizeof(Cplx<R>) * L, hipMemcpyHostToDevice) != hipSuccess) return kHipError;
  if (half_angles) {
    for (int i = 0; i < L; i++) {
      // angle -pi i / L in FP64
      const double a = -3.14159265358979323846 * (double)i / (double)L;
      h[i] = cmk<R>(R(cos(a)), R(sin(a)));
    }
    if (hipMalloc(&tw2, sizeof(Cplx<R>) * L) != hipSuccess) return kHipError;
    if (hipMemcpy(tw2, h.data(), sizeof(Cplx<R>) * L, hipMemcpyHostToDevice) != hipSuccess) return kHipError;
  }
  return kOk;
}

template <class R>
static void launch_pass(int kind, dim3 grid, const PassArgs<R>& P, hipStream_t s) {
  (void)with_length<R>(P.L, [&](auto Lc) {
    pass_launch<R, decltype(Lc)::value>(kind, grid, P, s);
    return kOk;
  });
}

// per-axis factors of the rows pass: complex ones (ctab, interleaved pairs)
// take precedence over the real D factors (tab)
template <class R>
static void set_factors(PassArgs<R>& P, const double* tab, const double* ctab, int N0, int N1) {
  if (ctab) {
    P.tab0 = ctab;
    P.tab1 = ctab + 2 * N0;
    P.tab2 = ctab + 2 * (N0 + N1);
    P.ctab = 1;
  } else if (tab) {
    P.tab0 = tab;
    P.tab1 = tab + N0;
    P.tab2 = tab + N0 + N1;
  }
  P.N1 = N1;
}

template <class R>
int fft3d_forward(const Fft3d<R>& F, const Cplx<R>* fhat, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                  Cplx<R>* grid, hipStream_t s, bool half, hipEvent_t mark1, hipEvent_t mark2,
                  const double* ctab) {
  const int N0 = F.N[0], N1 = F.N[1], N2 = F.N[2];
  const int n0 = F.n[0], n1 = F.n[1], n2 = F.n[2];
  if (half != F.c2r) return kUnsupported;
  if (half) {
    const int H = N2 / 2 + 1;
    // axis 1: f_hat[i0][i1][k2] (D fused on the loads) -> buf1[i0][p1][k2]
    {
      PassArgs<R> P = make_args(F.ax[1]);
      P.in = fhat; P.out = buf1;
      P.in_outer = (long)N1 * H; P.in_elem = H;
      P.out_outer = (long)n1 * H; P.out_elem = H;
      P.n_inner = H;
      P.C = F.ax[1].Ccols_narrow;
      P.Lin = N1; P.Lout = n1;
      P.h = N1 / 2; P.pos = N1 - N1 / 2;
      P.colfac = 1;
      set_factors(P, tab, ctab, N0, N1);
      dim3 grid_dim((H + P.C - 1) / P.C, N0);
      launch_pass<R>(kPassCols, grid_dim, P, s);
      if (mark1) (void)hipEventRecord(mark1, s);
    }
    // axis 0: pencils (p1, k2) stride n1*H -> buf2[p0][p1][k2]
    {
      PassArgs<R> P = make_args(F.ax[0]);
      P.in = buf1; P.out = buf2;
      P.in_outer = 0; P.in_elem = (long)n1 * H;
      P.out_outer = 0; P.out_elem = (long)n1 * H;
      P.n_inner = n1 * H;
      P.C = F.ax[0].Ccols_narrow;
      P.Lin = N0; P.Lout = n0;
      P.h = N0 / 2; P.pos = N0 - N0 / 2;
      dim3 grid_dim((n1 * H + P.C - 1) / P.C, 1);
      launch_pass<R>(kPassCols, grid_dim, P, s);
      if (mark2) (void)hipEventRecord(mark2, s);
    }
    // axis 2: c2r rows buf2[p0][p1][:] -> the real grid rows (n2 reals = n2/2 complex)
    {
      PassArgs<R> P = make_args(F.ax[2]);
      P.in = buf2; P.out = grid;
      P.in_outer = H; P.out_outer = n2 / 2;
      P.n_inner = n0 * n1;
      P.C = FftEpt<R>::value * kThreads / F.ax[2].L;
      P.tw2 = F.ax[2].tw2;
      P.N2h = N2 / 2;
      dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
      launch_pass<R>(kPassRowsC2r, grid_dim, P, s);
    }
    return hipGetLastError() == hipSuccess ? kOk : kHipError;
  }
  // axis 2: rows (i0,i1) of f_hat -> buf1[i0][i1][p2]
  {
    PassArgs<R> P = make_args(F.ax[2]);
    P.in = fhat; P.out = buf1;
    P.in_outer = N2; P.in_elem = 1;
    P.out_outer = n2; P.out_elem = 1;
    P.n_inner = N0 * N1;
    P.Lin = N2; P.Lout = n2;
    P.h = N2 / 2;
    P.pos = N2 - N2 / 2;
    set_factors(P, tab, ctab, N0, N1);
    dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
    launch_pass<R>(kPassRows, grid_dim, P, s);
    if (mark1) (void)hipEventRecord(mark1, s);
  }
  // axis 1: pencils (i0, p2) stride n2 -> buf2[i0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[1]);
    P.in = buf1; P.out = buf2;
    P.in_outer = (long)N1 * n2; P.in_elem = n2;
    P.out_outer = (long)n1 * n2; P.out_elem = n2;
    P.n_inner = n2;
    P.C = F.ax[1].Ccols_narrow;
    P.Lin = N1; P.Lout = n1;
    P.h = N1 / 2; P.pos = N1 - N1 / 2;
    dim3 grid_dim((n2 + P.C - 1) / P.C, N0);
    launch_pass<R>(kPassCols, grid_dim, P, s);
    if (mark2) (void)hipEventRecord(mark2, s);
  }
  // axis 0: pencils (p1, p2) stride n1*n2 -> grid[p0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[0]);
    P.in = buf2; P.out = grid;
    P.in_outer = n2; P.in_elem = (long)n1 * n2;
    P.out_outer = n2; P.out_elem = (long)n1 * n2;
    P.n_inner = n2;
    P.C = F.ax[0].Ccols;
    P.Lin = N0; P.Lout = n0;
    P.h = N0 / 2; P.pos = N0 - N0 / 2;
    dim3 grid_dim((n2 + P.C - 1) / P.C, n1);
    launch_pass<R>(kPassCols, grid_dim, P, s);
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int fft3d_backward(const Fft3d<R>& F, const Cplx<R>* grid, const double* tab, Cplx<R>* buf1, Cplx<R>* buf2,
                   Cplx<R>* fhat, int accumulate, hipStream_t s, bool half, hipEvent_t mark1,
                   hipEvent_t mark2, const double* ctab) {
  const int N0 = F.N[0], N1 = F.N[1], N2 = F.N[2];
  const int n0 = F.n[0], n1 = F.n[1], n2 = F.n[2];
  if (half != F.c2r) return kUnsupported;
  if (half) {
    const int H = N2 / 2 + 1;
    // axis 2: r2c rows of the real grid -> buf2[p0][p1][k2], k2 in [-N2/2, 0]
    {
      PassArgs<R> P = make_args(F.ax[2]);
      P.in = grid; P.out = buf2;
      P.in_outer = n2 / 2; P.out_outer = H;
      P.n_inner = n0 * n1;
      P.C = FftEpt<R>::value * kThreads / F.ax[2].L;
      P.tw2 = F.ax[2].tw2;
      P.N2h = N2 / 2;
      P.backward = 1;
      dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
      launch_pass<R>(kPassRowsR2c, grid_dim, P, s);
      if (mark1) (void)hipEventRecord(mark1, s);
    }
    // axis 0: buf2 pencils (p1, k2) -> buf1[i0][p1][k2]
    {
      PassArgs<R> P = make_args(F.ax[0]);
      P.in = buf2; P.out = buf1;
      P.in_outer = 0; P.in_elem = (long)n1 * H;
      P.out_outer = 0; P.out_elem = (long)n1 * H;
      P.n_inner = n1 * H;
      P.C = F.ax[0].Ccols_narrow;
      P.Lin = n0; P.Lout = N0; P.h = N0 / 2;
      P.backward = 1;
      dim3 grid_dim((n1 * H + P.C - 1) / P.C, 1);
      launch_pass<R>(kPassCols, grid_dim, P, s);
      if (mark2) (void)hipEventRecord(mark2, s);
    }
    // axis 1: buf1[i0][p1][k2] -> f_hat[i0][i1][k2] (D^T fused on the stores, accumulated)
    {
      PassArgs<R> P = make_args(F.ax[1]);
      P.in = buf1; P.out = fhat;
      P.in_outer = (long)n1 * H; P.in_elem = H;
      P.out_outer = (long)N1 * H; P.out_elem = H;
      P.n_inner = H;
      P.C = F.ax[1].Ccols_narrow;
      P.Lin = n1; P.Lout = N1; P.h = N1 / 2;
      P.backward = 1;
      P.accumulate = accumulate;
      P.colfac = 1;
      set_factors(P, tab, ctab, N0, N1);
      dim3 grid_dim((H + P.C - 1) / P.C, N0);
      launch_pass<R>(kPassCols, grid_dim, P, s);
    }
    return hipGetLastError() == hipSuccess ? kOk : kHipError;
  }
  const int N2out = N2;
  // axis 0: grid pencils (p1,p2) -> buf2[i0][p1][p2]
  {
    PassArgs<R> P = make_args(F.ax[0]);
    P.in = grid; P.out = buf2;
    P.in_outer = n2; P.in_elem = (long)n1 * n2;
    P.out_outer = n2; P.out_elem = (long)n1 * n2;
    P.n_inner = n2;
    P.C = F.ax[0].Ccols;
    P.Lin = n0; P.Lout = N0; P.h = N0 / 2;
    P.backward = 1;
    dim3 grid_dim((n2 + P.C - 1) / P.C, n1);
    launch_pass<R>(kPassCols, grid_dim, P, s);
    if (mark1) (void)hipEventRecord(mark1, s);
  }
  // axis 1: buf2[i0][p1][p2] -> buf1[i0][i1][p2]
  {
    PassArgs<R> P = make_args(F.ax[1]);
    P.in = buf2; P.out = buf1;
    P.in_outer = (long)n1 * n2; P.in_elem = n2;
    P.out_outer = (long)N1 * n2; P.out_elem = n2;
    P.n_inner = n2;
    P.C = F.ax[1].Ccols_narrow;
    P.Lin = n1; P.Lout = N1; P.h = N1 / 2;
    P.backward = 1;
    dim3 grid_dim((n2 + P.C - 1) / P.C, N0);
    launch_pass<R>(kPassCols, grid_dim, P, s);
    if (mark2) (void)hipEventRecord(mark2, s);
  }
  // axis 2: rows buf1[i0][i1][:] -> f_hat[i0][i1][:] (D^T fused)
  {
    PassArgs<R> P = make_args(F.ax[2]);
    P.in = buf1; P.out = fhat;
    P.in_outer = n2; P.in_elem = 1;
    P.out_outer = N2out; P.out_elem = 1;
    P.n_inner = N0 * N1;
    P.Lin = n2; P.Lout = N2out; P.h = N2 / 2;
    P.backward = 1;
    P.accumulate = accumulate;
    set_factors(P, tab, ctab, N0, N1);
    dim3 grid_dim((P.n_inner + P.C - 1) / P.C, 1);
    launch_pass<R>(kPassRows, grid_dim, P, s);
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int Fft3d<R>::init(const int* N_, const int* n_, bool c2r_) {
  c2r = c2r_;
  for (int t = 0; t < 3; t++) {
    N[t] = N_[t];
    n[t] = n_[t];
    if (N[t] > n[t]) return kUnsupported;
    // c2r: the real axis-2 transform of length n2 is a complex one of length n2 / 2
    const bool real_axis = c2r && t == 2;
    if (real_axis && n[t] < 4) return kUnsupported;
    int st = ax[t].init(real_axis ? n[t] / 2 : n[t], real_axis);
    if (st != kOk) return st;
  }
  for (int t = 0; t < 3; t++) {
    const int st = with_length<R>(ax[t].L, [](auto Lc) { return pass_attr<R, decltype(Lc)::value>(); });
    if (st != kOk) return st;
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
                                  Cplx<float>*, Cplx<float>*, hipStream_t, bool, hipEvent_t, hipEvent_t,
                                  const double*);
template int fft3d_forward<double>(const Fft3d<double>&, const Cplx<double>*, const double*, Cplx<double>*,
                                   Cplx<double>*, Cplx<double>*, hipStream_t, bool, hipEvent_t, hipEvent_t,
                                   const double*);
template int fft3d_backward<float>(const Fft3d<float>&, const Cplx<float>*, const double*, Cplx<float>*,
                                   Cplx<float>*, Cplx<float>*, int, hipStream_t, bool, hipEvent_t, hipEvent_t,
                                   const double*);
template int fft3d_backward<double>(const Fft3d<double>&, const Cplx<double>*, const double*, Cplx<double>*,
                                    Cplx<double>*, Cplx<double>*, int, hipStream_t, bool, hipEvent_t, hipEvent_t,
                                    const double*);

}  // namespace pnfft_amd
