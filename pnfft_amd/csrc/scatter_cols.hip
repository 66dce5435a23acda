// scatter_cols.hip -- adjoint window scatter B^T with register accumulators.
//
// This file is compiled once per (precision, mode); the build passes
// PNFFT_B_REAL (float|double) and PNFFT_B_MODE (B_F, B_GRAD or B_FGRAD).
//
// Reference: the scatter node loop of kernel/ndft-parallel.c:2890-3009 with
// the stencil kernels spread_f_c2c / spread_grad_f_c2c of
// kernel/assign.c:478-640 (cutoff 2m+1, kernel/ndft-parallel.c:910).
//
// One block per nonempty tile, tiles with more nodes first.  The tile's padded
// subgrid (P = TS + 2m cells per axis from the tile origin; every node binned to
// the tile has its stencil inside it) lives in registers: thread (b, c) owns
// the column (., b, c), and for FP32 f with m <= 4 also (., b + P/2, c), and keeps
// their cells as FP64 (re, im) accumulators.
// The tile's nodes are taken in sorted order, in batches: one thread per node
// evaluates the window values into LDS, then each node is applied by the
// threads whose columns lie in its stencil; the node's x-window is stored
// shifted to its plane offset in a zeroed row of P planes, so every thread
// updates its accumulators with fixed register indices (two-plane blocks the
// window does not reach are skipped by a uniform branch).
// Each thread then writes its column to the tile's slot, and the fold kernel
// sums, for every grid cell, the slots of the tiles whose subgrids cover it in
// a fixed order and writes the cell.  Accumulation and fold use no atomics, so
// the adjoint is deterministic; only nodes whose bins are older than their
// coordinates (listed by the column kernel) go through the direct path
// (global atomics) afterwards.
#include "b_device.hpp"

namespace pnfft_amd {

namespace {

using namespace bdev;

template <class R, int MT, int MODE>
struct ColsShape {
  static constexpr int CM = 2 * MT + 1;
  static constexpr int TS = std_tile(MT, sizeof(R) == 8);
  static constexpr int P = TS + 2 * MT;  // even
  static constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  static constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  // columns per thread: two (rows b and b + P/2) where their FP64 accumulators
  // fit in 256 VGPRs with the rest of the kernel (FP32, f only, m <= 4), else one
  static constexpr int NC = (sizeof(R) == 4 && MT <= 4 && !WG) ? 2 : 1;
  static constexpr int H = P / NC;       // rows covered by the first column of each thread
  static constexpr int NCOL = H * P;     // threads that own columns
  static constexpr int NT = (NCOL + 63) / 64 * 64;
  static constexpr int NB = WG ? 64 : 128;  // nodes per batch
};

// Planes are updated in blocks of two; block k (planes 2k, 2k+1) is skipped by
// a uniform branch when the node's window [l0, l0 + CM) does not reach it.
template <int NC, int P, int CM, bool WG, int K>
__device__ __attribute__((always_inline)) inline void plane_blocks(int l0, double (&ar)[NC][P], double (&ai)[NC][P],
                                                                   const double* __restrict__ px,
                                                                   const double* __restrict__ dx,
                                                                   const double (&w)[NC][4]) {
  if constexpr (2 * K < P) {
    if (2 * K + 1 >= l0 && 2 * K < l0 + CM) {
#pragma unroll
      for (int a = 2 * K; a < 2 * K + 2; a++) {
        const double x = px[a];
        const double d = WG ? dx[a] : 0.0;
#pragma unroll
        for (int r = 0; r < NC; r++) {
          ar[r][a] = fma(x, w[r][0], ar[r][a]);
          ai[r][a] = fma(x, w[r][1], ai[r][a]);
          if (WG) {
            ar[r][a] = fma(d, w[r][2], ar[r][a]);
            ai[r][a] = fma(d, w[r][3], ai[r][a]);
          }
        }
      }
    }
    plane_blocks<NC, P, CM, WG, K + 1>(l0, ar, ai, px, dx, w);
  }
}

// Weights of node i for the thread's rows r (b, b + H): w[r] = (re, im) of the
// f term and of the d/dx term, zero outside the stencil.  Returns false when
// the thread's column is outside the node's stencil (or the node is outside
// the tile, loc < 0).
template <class R, int NC, int CM, bool WF, bool WG, int NB, int NBG>
__device__ __attribute__((always_inline)) inline bool node_weights(int loc, int i, int b, int c, int H, bool own,
                                                                   const R (&s_py)[NB][CM], const R (&s_pz)[NB][CM],
                                                                   const R (&s_dy)[NBG][CM], const R (&s_dz)[NBG][CM],
                                                                   const Cplx<R> (&s_f)[WF ? NB : 1],
                                                                   const Cplx<R> (&s_g)[NBG][3], double (&w)[NC][4]) {
  if (loc < 0) return false;
  const int l1 = (loc >> 8) & 255, l2 = (loc >> 16) & 255;
  const int dc = c - l2;
  if (!own || (unsigned)dc >= (unsigned)CM) return false;
  const R pz = s_pz[i][dc];
  const R dz = WG ? s_dz[i][dc] : R(0);
#pragma unroll
  for (int r = 0; r < NC; r++) {
    const int db = b + r * H - l1;
    const bool in_b = (unsigned)db < (unsigned)CM;
    const int dbc = in_b ? db : 0;
    const R py = in_b ? s_py[i][dbc] : R(0);
    w[r][0] = w[r][1] = w[r][2] = w[r][3] = 0.0;
    if (WF) {
      const Cplx<R> f = s_f[i];
      const R q = py * pz;
      w[r][0] = (double)(q * f.x);
      w[r][1] = (double)(q * f.y);
    }
    if (WG) {
      const R dy = in_b ? s_dy[i][dbc] : R(0);
      const Cplx<R> g0 = s_g[i][0], g1 = s_g[i][1], g2 = s_g[i][2];
      const R q0 = py * pz, q1 = dy * pz, q2 = py * dz;
      w[r][2] = (double)(q0 * g0.x);
      w[r][3] = (double)(q0 * g0.y);
      w[r][0] += (double)(q1 * g1.x + q2 * g2.x);
      w[r][1] += (double)(q1 * g1.y + q2 * g2.y);
    }
  }
  return true;
}

template <class R, int MT, int MODE>
__global__ __launch_bounds__((ColsShape<R, MT, MODE>::NT)) void scatter_cols_k(BArgs<R> A) {
  using S = ColsShape<R, MT, MODE>;
  constexpr int CM = S::CM, P = S::P, H = S::H, NB = S::NB, TS = S::TS, NC = S::NC;
  constexpr bool WF = S::WF, WG = S::WG;
  __shared__ int s_loc[NB];
  __shared__ Cplx<R> s_f[WF ? NB : 1];
  __shared__ Cplx<R> s_g[WG ? NB : 1][3];
  // x-windows placed at the node's plane offset in a zeroed row of P planes, so
  // that every accumulator has a fixed register (an index l0 + a would not)
  __shared__ __align__(16) double s_px[NB][P];
  __shared__ __align__(16) double s_dx[WG ? NB : 1][P];
  __shared__ R s_py[NB][CM], s_pz[NB][CM];
  __shared__ R s_dy[WG ? NB : 1][CM], s_dz[WG ? NB : 1][CM];

  const int tile = A.torder[blockIdx.x];
  const int t2 = tile % A.ntile[2];
  const int t1 = (tile / A.ntile[2]) % A.ntile[1];
  const int t0 = tile / (A.ntile[2] * A.ntile[1]);
  const int o0 = t0 * TS, o1 = t1 * TS, o2 = t2 * TS;
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const int beg = A.tile_start[tile], end = A.tile_start[tile + 1];
  const int tid = threadIdx.x;
  const int b = tid / P, c = tid % P;  // rows b (and b + H when NC = 2) of column c
  const bool own = tid < S::NCOL;

  double ar[NC][P], ai[NC][P];
#pragma unroll
  for (int r = 0; r < NC; r++)
#pragma unroll
    for (int a = 0; a < P; a++) ar[r][a] = ai[r][a] = 0.0;

  for (int base = beg; base < end; base += NB) {
    const int nb = min(NB, end - base);
    if (tid < nb) {
      const long j = A.perm[base + tid];
      NodeWin<R, CM> w;
      node_windows<R, MT, CM, WG>(A, base + tid, j, MT, w);
      const int l0 = local_off(w.start[0], o0, TS, n0);
      const int l1 = local_off(w.start[1], o1, TS, n1);
      const int l2 = local_off(w.start[2], o2, TS, n2);
      const bool in = l0 >= 0 && l1 >= 0 && l2 >= 0;
      s_loc[tid] = in ? (l0 | (l1 << 8) | (l2 << 16)) : -1;
      if (!in) A.stale_list[atomicAdd(A.stale_count, 1u)] = base + tid;
#pragma unroll
      for (int a = 0; a < P; a++) {
        s_px[tid][a] = 0.0;
        if (WG) s_dx[tid][a] = 0.0;
      }
      const int o = in ? l0 : 0;
#pragma unroll
      for (int a = 0; a < CM; a++) {
        s_px[tid][o + a] = (double)w.px[a];
        s_py[tid][a] = w.py[a];
        s_pz[tid][a] = w.pz[a];
        if (WG) {
          s_dx[tid][o + a] = (double)w.dx[a];
          s_dy[tid][a] = w.dy[a];
          s_dz[tid][a] = w.dz[a];
        }
      }
      if (WF) s_f[tid] = load_f(A, j);
      if (WG) {
        s_g[tid][0] = load_g(A, j, 0);
        s_g[tid][1] = load_g(A, j, 1);
        s_g[tid][2] = load_g(A, j, 2);
      }
    }
    __syncthreads();
    for (int i = 0; i < nb; i++) {
      const int loc = __builtin_amdgcn_readfirstlane(s_loc[i]);
      double w[NC][4];
      if (node_weights<R, NC, CM, WF, WG>(loc, i, b, c, H, own, s_py, s_pz, s_dy, s_dz, s_f, s_g, w))
        plane_blocks<NC, P, CM, WG, 0>(loc & 255, ar, ai, &s_px[i][0], WG ? &s_dx[i][0] : nullptr, w);
    }
    __syncthreads();
  }

  if (own) {
    Cplx<R>* slot = A.slots + (size_t)blockIdx.x * (size_t)(P * P * P);
#pragma unroll
    for (int r = 0; r < NC; r++)
#pragma unroll
      for (int a = 0; a < P; a++) slot[(a * P + b + r * H) * P + c] = cmk<R>((R)ar[r][a], (R)ai[r][a]);
  }
}

// ---------------------------------------------------------------------------
// Matrix-core form (FP32 plans, P = TS + 2m <= 16, i.e. m <= 6): the tile update
//   G[a][b][c] += sum_v X_v[a] Y_v[b] W_v Z_v[c]
// over rank-one terms v (one per node for f; three for grad, (psi'x, psi y,
// psi z; g0), (psi x, psi'y, psi z; g1), (psi x, psi y, psi'z; g2); four for
// f + grad) is, for every plane a, the product of the 16 x K matrix
// A[b][v] = X_v[a] Y_v[b] and the K x 16 matrix B[v][c] = W_v Z_v[c]: one
// v_mfma_f32_16x16x4f32 per plane, K = 4 terms, per real and imaginary part.
// The four waves of a block own the planes a = w + 4q (q = 0..3, interleaved so
// that every wave meets a node's stencil on 2-3 of its planes); a plane is
// skipped for a whole batch when none of the batch's stencils reaches it
// (sorted nodes of a tile share their plane offset most of the time).  The
// FP32 accumulators sum over all nodes of the tile and go to the tile's slot.
// ---------------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template <int MODE>
struct MfmaShape {
  static constexpr int V = MODE == B_F ? 1 : (MODE == B_GRAD ? 3 : 4);  // rank-one terms per node
  static constexpr int NB = 64;                                          // nodes per batch
  static constexpr int NV = NB * V;                                      // terms per batch
  // waves per SIMD the kernel is compiled for: the gradient windows need the larger register budget
  static constexpr int WAVES = V == 1 ? 3 : 2;
  // row buffers: two for f (rows of the next batch written during this batch's
  // matrix-core phase); the gradient modes' three or four terms per node keep one
  static constexpr int NBUF = V == 1 ? 2 : 1;
};

// Window rows of the batch: term v's row of axis t holds the window at the
// cells l .. l + 2m of the node's stencil (zero elsewhere) and starts at
// 16 v + v / 2.  The half-word step per row pair keeps both access patterns in
// distinct banks: the rows phase, where lane q writes the same position of row
// q (V = 1), and the matrix-core phase, where 32 lanes read two consecutive
// rows (v even, v + 1).
__device__ __forceinline__ constexpr int row_at(int v) { return 16 * v + (v >> 1); }

// OR of v over each row of 16 lanes (DPP row shifts); lane 15 of a row holds it
__device__ __forceinline__ unsigned row_or16(unsigned v) {
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  return v;
}

// OR of v over the 64 lanes of a wave, as a wave-uniform value
__device__ __forceinline__ unsigned wave_or(unsigned v) {
  v = row_or16(v);
  return (unsigned)(__builtin_amdgcn_readlane((int)v, 15) | __builtin_amdgcn_readlane((int)v, 31) |
                    __builtin_amdgcn_readlane((int)v, 47) | __builtin_amdgcn_readlane((int)v, 63));
}

// one thread's share of a node's data for the rows phase: axis `role` (0..2)
// or the complex weights (role 3)
template <int CM, bool WF, bool WG>
struct ScatterLoad {
  int j;     // node id, -1: no node
  float x;   // coordinate now ...
  float px;  // ... and when the tables were made
  float psi[CM];
  float dpsi[WG ? CM : 1];
  Cplx<float> f;
  Cplx<float> g[WG ? 3 : 1];
};

// RG: the real grid of a c2r plan (real node values: the imaginary products are left out)
template <int MT, int MODE, bool RG>
__global__ __launch_bounds__(256, (MfmaShape<MODE>::WAVES + (MODE == B_F && MT <= 4 ? 1 : 0))) void scatter_mfma_k(
    BArgs<float> A) {
  using S = ColsShape<float, MT, MODE>;
  using MS = MfmaShape<MODE>;
  constexpr int CM = S::CM, P = S::P, TS = S::TS, V = MS::V, NB = MS::NB, NV = MS::NV;
  constexpr bool WF = S::WF, WG = S::WG;
  static_assert(P <= 16, "one 16 x 16 block per plane");
  static_assert(NB == 64, "one node per lane of a wave in the rows phase");
  // window rows of a batch; with two buffers the rows of batch b + 1 are written
  // while batch b is on the matrix cores (one barrier per batch)
  constexpr int NBUF = MS::NBUF;
  constexpr int RS = row_at(NV);                   // floats per row buffer
  __shared__ __align__(16) float s_x[NBUF][RS];    // X_v[a] at row_at(v) + a, zero outside the stencil
  __shared__ __align__(16) float s_y[NBUF][RS];    // Y_v[b]
  __shared__ __align__(16) float s_z[NBUF][RS];    // Z_v[c]
  __shared__ float2 s_w[NBUF][NV];                 // W_v
  __shared__ unsigned char s_bad[NBUF][3][NB];     // node outside the tile on axis t
  __shared__ unsigned s_pm[NBUF];                  // planes the batch's stencils reach (bit a)
  static_assert(RS % 4 == 0, "row buffers of whole 16-byte chunks");

  const int tile = A.torder[blockIdx.x];
  const int t2 = tile % A.ntile[2];
  const int t1 = (tile / A.ntile[2]) % A.ntile[1];
  const int t0 = tile / (A.ntile[2] * A.ntile[1]);
  const int o0 = t0 * TS, o1 = t1 * TS, o2 = t2 * TS;
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const int beg = A.tile_start[tile], end = A.tile_start[tile + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  // rows phase: thread (role, q) handles node q of the batch; roles 0..2 the
  // window of axis `role`, role 3 the weights
  const int q = lane, role = wave;
  const int ax = role < 3 ? role : 0;
  const int nt = ax == 0 ? n0 : (ax == 1 ? n1 : n2);
  const int ot = ax == 0 ? o0 : (ax == 1 ? o1 : o2);
  const float bt = (float)(ax == 0 ? A.g.b[0] : (ax == 1 ? A.g.b[1] : A.g.b[2]));
  const float wst = (float)(ax == 0 ? A.g.ws[0] : (ax == 1 ? A.g.ws[1] : A.g.ws[2]));
  const long M = A.M;
  const bool tab = A.pre_psi != nullptr && (!WG || A.pre_dpsi != nullptr);
  // With box slots the slot holds only the box the tile's nodes reached at
  // binning time (per axis, the smallest offset to the largest plus 2m); a node
  // whose offset now lies outside it takes the stale path like a node outside
  // its tile.  bml: the offsets of this thread's axis.
  const unsigned bm = A.box_slots ? A.box_ord[blockIdx.x] : 0x00FFFFFFu;
  const unsigned bml = (unsigned)__builtin_amdgcn_readfirstlane((int)((bm >> (8 * ax)) & 0xFFu));  // one axis per wave
  const int blo = bml ? __builtin_ctz(bml) : 0, bhi = bml ? 31 - __builtin_clz(bml) : -1;

  // The loads below are unconditional, from clamped positions (their values
  // are discarded where no node exists), so that no register copy of a loaded
  // value is needed before the matrix-core phase; such a copy would make the
  // wave wait for every load in flight.
  auto perm_raw = [&](int base) { return A.perm[min(base + q, end - 1)]; };
  auto perm_ok = [&](int base) { return base + q < end; };
  auto load = [&](int base, int j, ScatterLoad<CM, WF, WG>& L) {
    L.j = j;
    const int jj = max(j, 0);
    if (role < 3) {
      L.x = A.x[3 * (long)jj + ax];
      if (tab) {
        const long p = min(base + q, end - 1);
        L.px = A.pre_x[(long)ax * M + p];
#pragma unroll
        for (int s = 0; s < CM; s++) L.psi[s] = A.pre_psi[(long)(ax * CM + s) * M + p];
        if (WG)
#pragma unroll
          for (int s = 0; s < CM; s++) L.dpsi[s] = A.pre_dpsi[(long)(ax * CM + s) * M + p];
      }
    } else {
      if (WF) L.f = load_f(A, jj);
      if (WG)
#pragma unroll
        for (int t = 0; t < 3; t++) L.g[t] = load_g(A, jj, t);
    }
  };

  // plane wave + 4q: G[b = 4 lk + e][c = li], summed in FP32 over the tile's
  // nodes (the single-precision reference sums its adjoint in float as well)
  mfma_f4 cre[4], cim[4];
#pragma unroll
  for (int qq = 0; qq < 4; qq++)
#pragma unroll
    for (int e = 0; e < 4; e++) cre[qq][e] = cim[qq][e] = 0.f;

  ScatterLoad<CM, WF, WG> L;
  {
    const int j0 = perm_raw(beg);
    load(beg, perm_ok(beg) ? j0 : -1, L);
  }
  int jn = perm_raw(beg + NB < end ? beg + NB : beg);
  bool jn_ok = perm_ok(beg + NB);

  // rows phase: window rows (roles 0..2) and weights (role 3) of the batch at
  // base, from the loads in L, into row buffer bf
  auto rows_phase = [&](const ScatterLoad<CM, WF, WG>& L, int bf) __attribute__((always_inline)) {
    if (role < 3) {
      float w[CM], dw[CM];
      int l = 0;
      bool bad = false;
      if (L.j >= 0) {
        const NodeAxis<float> na = node_axis<float>(L.x, nt, MT, A.g.shift);
        l = local_off(na.start, ot, TS, nt);
        bad = l < blo || l > bhi;  // outside the tile (l < 0) or outside the slot's box
        if (tab && L.x == L.px) {
#pragma unroll
          for (int s = 0; s < CM; s++) {
            w[s] = L.psi[s];
            dw[s] = WG ? L.dpsi[s] : 0.f;
          }
        } else {
          // inlined: an out-of-line call would keep w and dw in scratch memory
          [[clang::always_inline]] window_axis_s<float, CM, WG>(A.g.win, na.frac, (float)nt, bt, w, dw);
#pragma unroll
          for (int s = 0; s < CM; s++) {
            w[s] *= wst;
            if (WG) dw[s] *= wst;
          }
        }
      }
      float* srow = ax == 0 ? s_x[bf] : (ax == 1 ? s_y[bf] : s_z[bf]);
      // the wave's rows are zeroed with consecutive lanes on consecutive 16-byte
      // chunks; the window values below are later stores of the same wave
      const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < (RS / 4 + 63) / 64; k++)
        if (q + 64 * k < RS / 4) reinterpret_cast<float4*>(srow)[q + 64 * k] = z4;
#pragma unroll
      for (int u = 0; u < V; u++) {
        float* row = srow + row_at(q * V + u);
        if (L.j >= 0 && !bad) {
          // term u differentiates axis ax when it is the gradient term of that axis
          const bool d = WG && u == (WF ? 1 : 0) + ax;
#pragma unroll
          for (int s = 0; s < CM; s++) row[l + s] = d ? dw[s] : w[s];
        }
      }
      s_bad[bf][ax][q] = bad ? 1 : 0;
      if (role == 0) {
        // the planes of the batch: the union of its nodes' x stencils
        const unsigned pm = wave_or((L.j >= 0 && !bad) ? (((1u << CM) - 1u) << l) : 0u);
        if (q == 0) s_pm[bf] = pm;
      }
    } else {
      int v = q * V;
      if (WF) s_w[bf][v++] = L.j >= 0 ? make_float2(L.f.x, L.f.y) : make_float2(0.f, 0.f);
      if (WG)
#pragma unroll
        for (int t = 0; t < 3; t++) s_w[bf][v++] = L.j >= 0 ? make_float2(L.g[t].x, L.g[t].y) : make_float2(0.f, 0.f);
    }
  };

  if (NBUF == 2) {  // rows of the first batch
    rows_phase(L, 0);
    ScatterLoad<CM, WF, WG> Ln;
    const int bn = beg + NB < end ? beg + NB : beg;
    load(bn, jn_ok ? jn : -1, Ln);
    jn = perm_raw(beg + 2 * NB < end ? beg + 2 * NB : beg);
    jn_ok = perm_ok(beg + 2 * NB);
    L = Ln;
    __syncthreads();
  }
  for (int base = beg, it = 0; base < end; base += NB, it++) {
    const int nb = min(NB, end - base);
    const int bf = NBUF == 2 ? (it & 1) : 0;
    if (NBUF == 1) {
      rows_phase(L, 0);
      __syncthreads();
    }
    // a node outside its tile has a zero row on the axis where it left, so it
    // adds nothing here; it is listed for the direct kernel
    if (role == 3 && base + q < end && (s_bad[bf][0][q] | s_bad[bf][1][q] | s_bad[bf][2][q]))
      A.stale_list[atomicAdd(A.stale_count, 1u)] = base + q;
    if (NBUF == 2 && base + NB < end) rows_phase(L, bf ^ 1);  // the next batch, into the other buffer
    // loads of the batch after the one in L, in flight during the matrix-core phase
    ScatterLoad<CM, WF, WG> Ln;
    const int bnext = base + NBUF * NB;
    load(bnext < end ? bnext : base, jn_ok ? jn : -1, Ln);
    const int jnn = perm_raw(bnext + NB < end ? bnext + NB : base);
    const bool jnn_ok = perm_ok(bnext + NB);
    const int nq = (nb * V + 3) >> 2;
    // this wave's planes a = wave + 4 qq that the batch reaches (bit qq)
    const unsigned pmb = (unsigned)__builtin_amdgcn_readfirstlane((int)s_pm[bf]) >> wave;
    const unsigned act = (pmb & 1u) | ((pmb >> 3) & 2u) | ((pmb >> 6) & 4u) | ((pmb >> 9) & 8u);
    // term v = 4 q4 + lk of this lane group: row_at(v) = row_at(lk) + 66 q4
    const float* xrow = s_x[bf] + row_at(lk) + wave;
    const float* yrow = s_y[bf] + row_at(lk) + li;
    const float* zrow = s_z[bf] + row_at(lk) + li;
    const float2* wrow = &s_w[bf][lk];
#pragma unroll 4
    for (int q4 = 0; q4 < nq; q4++) {
      const float y = yrow[66 * q4], z = zrow[66 * q4];
      const float2 wv = wrow[4 * q4];
      const float bre = wv.x * z, bim = wv.y * z;
#pragma unroll
      for (int qq = 0; qq < 4; qq++) {
        if (act & (1u << qq)) {
          const float a = xrow[66 * q4 + 4 * qq] * y;
          cre[qq] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bre, cre[qq], 0, 0, 0);
          if constexpr (!RG) cim[qq] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bim, cim[qq], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    L = Ln;
    jn = jnn;
    jn_ok = jnn_ok;
  }

  Cplx<float>* slot = A.slots + (size_t)blockIdx.x * (size_t)(P * P * P);
  // cells of the box per axis (all cells without box slots)
  auto cells = [](unsigned m) {  // offsets m -> cells lo .. hi + 2m
    if (!m) return 0u;
    const int lo = __builtin_ctz(m), hi = 31 - __builtin_clz(m) + CM - 1;
    return ((2u << hi) - 1u) & ~((1u << lo) - 1u);
  };
  const unsigned cx = cells(bm & 0xFFu), cy = cells((bm >> 8) & 0xFFu), cz = cells((bm >> 16) & 0xFFu);
#pragma unroll
  for (int qq = 0; qq < 4; qq++) {
    const int a = wave + 4 * qq;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int b = 4 * lk + e;
      if (a < P && b < P && li < P && ((cx >> a) & (cy >> b) & (cz >> li) & 1u))
        slot[(a * P + b) * P + li] = cmk<float>(cre[qq][e], cim[qq][e]);
    }
  }
}

// grid[x] = sum over the tiles whose padded subgrids cover x of their slot
// value at x, tiles taken in a fixed order; cells no nonempty tile covers get 0.
// One block per tile of the grid (empty ones included), one thread per cell;
// the neighbour tile (T - d) is the same for the whole block, so empty
// neighbours are skipped by a uniform branch.
template <class R, int TS, int P>
__global__ __launch_bounds__(TS* TS* TS) void fold_k(BArgs<R> A) {
  constexpr int D = (P + TS - 1) / TS;  // tiles per axis whose subgrids can cover a cell
  constexpr size_t VOL = (size_t)P * P * P;
  const int nt0 = A.ntile[0], nt1 = A.ntile[1], nt2 = A.ntile[2];
  const int T = blockIdx.x;
  const int T2 = T % nt2, T1 = (T / nt2) % nt1, T0 = T / (nt2 * nt1);
  const int q2 = threadIdx.x % TS, q1 = (threadIdx.x / TS) % TS, q0 = threadIdx.x / (TS * TS);
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int d0 = 0; d0 < D; d0++) {
    const int s0 = (T0 - d0 % nt0 + nt0) % nt0;
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      const int s1 = (T1 - d1 % nt1 + nt1) % nt1;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) {
        const int s2 = (T2 - d2 % nt2 + nt2) % nt2;
        const int slot = A.tile_slot[((long)s0 * nt1 + s1) * nt2 + s2];
        if (slot < 0) continue;
        const int p0 = q0 + d0 * TS, p1 = q1 + d1 * TS, p2 = q2 + d2 * TS;
        if (p0 < P && p1 < P && p2 < P) {
          const Cplx<R> v = A.slots[(size_t)slot * VOL + ((size_t)p0 * P + p1) * P + p2];
          sr += (double)v.x;
          si += (double)v.y;
        }
      }
    }
  }
  const long n1 = A.g.n[1], n2 = A.g.n[2];
  const long x0 = (long)T0 * TS + q0, x1 = (long)T1 * TS + q1, x2 = (long)T2 * TS + q2;
  if (A.c2r)  // real grid of a c2r plan
    reinterpret_cast<R*>(A.grid)[(x0 * n1 + x1) * n2 + x2] = (R)sr;
  else
    A.grid[(x0 * n1 + x1) * n2 + x2] = cmk<R>((R)sr, (R)si);
}

// fold_k for tiles of 8^3 cells: 128 threads per grid tile, four consecutive
// cells (along axis 2) per thread, loaded from each covering slot as two
// 16-byte loads where all four lie inside the subgrid; same sums, same order.
template <class R, int P>
__global__ __launch_bounds__(128) void fold8_k(BArgs<R> A) {
  constexpr int TS = 8;
  constexpr int D = (P + TS - 1) / TS;
  constexpr size_t VOL = (size_t)P * P * P;
  const int nt0 = A.ntile[0], nt1 = A.ntile[1], nt2 = A.ntile[2];
  const int T = blockIdx.x;
  const int T2 = T % nt2, T1 = (T / nt2) % nt1, T0 = T / (nt2 * nt1);
  const int h = threadIdx.x & 1, q1 = (threadIdx.x >> 1) & 7, q0 = threadIdx.x >> 4;
  const int q2 = 4 * h;  // first of the thread's four cells
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int d0 = 0; d0 < D; d0++) {
    const int s0 = (T0 - d0 % nt0 + nt0) % nt0;
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      const int s1 = (T1 - d1 % nt1 + nt1) % nt1;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) {
        const int s2 = (T2 - d2 % nt2 + nt2) % nt2;
        const int slot = A.tile_slot[((long)s0 * nt1 + s1) * nt2 + s2];
        if (slot < 0) continue;
        const int p0 = q0 + d0 * TS, p1 = q1 + d1 * TS, p2 = q2 + d2 * TS;
        if (p0 >= P || p1 >= P) continue;
        const Cplx<R>* src = A.slots + (size_t)slot * VOL + ((size_t)p0 * P + p1) * P + p2;
        if (p2 + 3 < P) {
          Cplx<R> v[4];
          if constexpr (sizeof(R) == 4) {  // two aligned 16-byte loads (p2 is a multiple of 4)
            const float4 a = reinterpret_cast<const float4*>(src)[0], b = reinterpret_cast<const float4*>(src)[1];
            v[0] = cmk<R>(a.x, a.y);
            v[1] = cmk<R>(a.z, a.w);
            v[2] = cmk<R>(b.x, b.y);
            v[3] = cmk<R>(b.z, b.w);
          } else {
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = src[k];
          }
#pragma unroll
          for (int k = 0; k < 4; k++) {
            sr[k] += (double)v[k].x;
            si[k] += (double)v[k].y;
          }
        } else {
#pragma unroll
          for (int k = 0; k < 4; k++) {
            if (p2 + k < P) {
              const Cplx<R> v = src[k];
              sr[k] += (double)v.x;
              si[k] += (double)v.y;
            }
          }
        }
      }
    }
  }
  const long n1 = A.g.n[1], n2 = A.g.n[2];
  const long x0 = (long)T0 * TS + q0, x1 = (long)T1 * TS + q1, x2 = (long)T2 * TS + q2;
  if (A.c2r) {  // real grid of a c2r plan
    R* dr = reinterpret_cast<R*>(A.grid) + (x0 * n1 + x1) * n2 + x2;
#pragma unroll
    for (int k = 0; k < 4; k++) dr[k] = (R)sr[k];
    return;
  }
  Cplx<R>* dst = A.grid + (x0 * n1 + x1) * n2 + x2;
  if constexpr (sizeof(R) == 4) {
    reinterpret_cast<float4*>(dst)[0] = make_float4((float)sr[0], (float)si[0], (float)sr[1], (float)si[1]);
    reinterpret_cast<float4*>(dst)[1] = make_float4((float)sr[2], (float)si[2], (float)sr[3], (float)si[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k] = cmk<R>((R)sr[k], (R)si[k]);
  }
}

// fold8_k for P <= 16 as bricks of two grid tiles along axis 2 (8 x 8 x 16
// cells, 256 threads, four consecutive cells along axis 2 per thread): a wave
// writes 128 contiguous bytes of a grid row (FP32).  Slot j = 0..2 along axis 2
// is tile T2 - 1 + j for the brick's first tile T2; cell c = 0..15 of the
// brick sits at p2 = c + 8 - 8j of it.  Every cell sums the same (slot, p2)
// pairs as fold8_k (two slots along axis 2, as nt2 >= 2), in the fixed order
// (d0, d1, j).
template <class R, int P>
__global__ __launch_bounds__(256) void fold8b_k(BArgs<R> A) {
  constexpr int TS = 8;
  constexpr int D = (P + TS - 1) / TS;
  static_assert(P % 4 == 0 && P <= 16, "aligned quads of cells inside a 16-cell subgrid");
  constexpr size_t VOL = (size_t)P * P * P;
  const int nt0 = A.ntile[0], nt1 = A.ntile[1], nt2 = A.ntile[2];
  const int nb2 = nt2 >> 1;  // bricks along axis 2
  const int B = blockIdx.x;
  const int T2 = (B % nb2) * 2, T1 = (B / nb2) % nt1, T0 = B / (nb2 * nt1);
  const int h = threadIdx.x & 3, q1 = (threadIdx.x >> 2) & 7, q0 = threadIdx.x >> 5;
  const int c = 4 * h;  // first of the thread's four cells in the brick
  auto cells = [](unsigned m) {  // offsets m -> cells lo .. hi + 2m of a slot
    if (!m) return 0u;
    const int lo = __builtin_ctz(m), hi = 31 - __builtin_clz(m) + P - TS;
    return ((2u << hi) - 1u) & ~((1u << lo) - 1u);
  };
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int d0 = 0; d0 < D; d0++) {
    const int s0 = (T0 - d0 % nt0 + nt0) % nt0;
    const int p0 = q0 + d0 * TS;
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      const int s1 = (T1 - d1 % nt1 + nt1) % nt1;
      const int p1 = q1 + d1 * TS;
      const bool row_ok = p0 < P && p1 < P;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int s2 = ((T2 - 1 + j) % nt2 + nt2) % nt2;
        const long tix = ((long)s0 * nt1 + s1) * nt2 + s2;
        const int slot = A.tile_slot[tix];
        if (slot < 0) continue;
        const int p2 = c + 8 - 8 * j;
        if (row_ok && p2 >= 0 && p2 < P) {
          // box slots: only the cells of the slot's box were written; zq marks
          // which of the four cells lie in it
          unsigned zq = 0xFu;
          if (A.box_slots) {
            const unsigned bmb = A.box_tile[tix];
            if (!((cells(bmb & 0xFFu) >> p0) & (cells((bmb >> 8) & 0xFFu) >> p1) & 1u)) continue;
            zq = (cells((bmb >> 16) & 0xFFu) >> p2) & 0xFu;
            if (!zq) continue;
          }
          const Cplx<R>* src = A.slots + (size_t)slot * VOL + ((size_t)p0 * P + p1) * P + p2;
          if constexpr (sizeof(R) == 4) {  // two aligned 16-byte loads (P and p2 multiples of 4)
            float4 a = reinterpret_cast<const float4*>(src)[0], b = reinterpret_cast<const float4*>(src)[1];
            if (zq != 0xFu) {
              if (!(zq & 1u)) a.x = a.y = 0.f;
              if (!(zq & 2u)) a.z = a.w = 0.f;
              if (!(zq & 4u)) b.x = b.y = 0.f;
              if (!(zq & 8u)) b.z = b.w = 0.f;
            }
            sr[0] += (double)a.x;
            si[0] += (double)a.y;
            sr[1] += (double)a.z;
            si[1] += (double)a.w;
            sr[2] += (double)b.x;
            si[2] += (double)b.y;
            sr[3] += (double)b.z;
            si[3] += (double)b.w;
          } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
              if (!((zq >> k) & 1u)) continue;
              sr[k] += (double)src[k].x;
              si[k] += (double)src[k].y;
            }
          }
        }
      }
    }
  }
  const long n1 = A.g.n[1], n2 = A.g.n[2];
  const long x0 = (long)T0 * TS + q0, x1 = (long)T1 * TS + q1, x2 = (long)T2 * TS + c;
  if (A.c2r) {  // real grid of a c2r plan
    R* dr = reinterpret_cast<R*>(A.grid) + (x0 * n1 + x1) * n2 + x2;
#pragma unroll
    for (int k = 0; k < 4; k++) dr[k] = (R)sr[k];
    return;
  }
  Cplx<R>* dst = A.grid + (x0 * n1 + x1) * n2 + x2;
  if constexpr (sizeof(R) == 4) {
    reinterpret_cast<float4*>(dst)[0] = make_float4((float)sr[0], (float)si[0], (float)sr[1], (float)si[1]);
    reinterpret_cast<float4*>(dst)[1] = make_float4((float)sr[2], (float)si[2], (float)sr[3], (float)si[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k] = cmk<R>((R)sr[k], (R)si[k]);
  }
}

// fold_k for tiles of 4^3 cells with P <= 16: one block of 128 threads per
// brick of four grid tiles along axis 2 (4 x 4 x 16 cells), two consecutive
// cells along axis 2 per thread.  For each neighbour offset (d0, d1) the seven
// slots j = 0..6 along axis 2 whose subgrids reach the brick (tile T2 - 3 + j
// for the brick's first tile T2) are read as whole 16-byte pairs: cell
// c = 0..15 of the brick sits at p2 = c + 12 - 4j of slot j, and a wave reads
// up to 128 contiguous bytes of a slot row.  Every cell sums the same
// (slot, p2) pairs as fold_k (four slots along axis 2, as nt2 >= 4), in the
// fixed order (d0, d1, j).
template <class R, int P>
__global__ __launch_bounds__(128) void fold4_k(BArgs<R> A) {
  constexpr int TS = 4;
  constexpr int D = (P + TS - 1) / TS;
  static_assert(P % 2 == 0 && P <= 16, "pairs of cells inside a 16-cell subgrid");
  constexpr size_t VOL = (size_t)P * P * P;
  const int nt0 = A.ntile[0], nt1 = A.ntile[1], nt2 = A.ntile[2];
  const int nb2 = nt2 >> 2;  // bricks along axis 2
  const int B = blockIdx.x;
  const int T2 = (B % nb2) * 4, T1 = (B / nb2) % nt1, T0 = B / (nb2 * nt1);
  const int h = threadIdx.x & 7, q1 = (threadIdx.x >> 3) & 3, q0 = threadIdx.x >> 5;
  const int c = 2 * h;  // first of the thread's two cells in the brick
  double sr[2] = {0.0, 0.0}, si[2] = {0.0, 0.0};
#pragma unroll
  for (int d0 = 0; d0 < D; d0++) {
    const int s0 = (T0 - d0 % nt0 + nt0) % nt0;
    const int p0 = q0 + d0 * TS;
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      const int s1 = (T1 - d1 % nt1 + nt1) % nt1;
      const int p1 = q1 + d1 * TS;
      const bool row_ok = p0 < P && p1 < P;
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const int s2 = ((T2 - 3 + j) % nt2 + nt2) % nt2;
        const int slot = A.tile_slot[((long)s0 * nt1 + s1) * nt2 + s2];
        if (slot < 0) continue;
        const int p2 = c + 12 - 4 * j;
        if (row_ok && p2 >= 0 && p2 < P) {
          const Cplx<R>* src = A.slots + (size_t)slot * VOL + ((size_t)p0 * P + p1) * P + p2;
          if constexpr (sizeof(R) == 4) {  // one aligned 16-byte load (P and p2 even)
            const float4 v = *reinterpret_cast<const float4*>(src);
            sr[0] += (double)v.x;
            si[0] += (double)v.y;
            sr[1] += (double)v.z;
            si[1] += (double)v.w;
          } else {
#pragma unroll
            for (int k = 0; k < 2; k++) {
              sr[k] += (double)src[k].x;
              si[k] += (double)src[k].y;
            }
          }
        }
      }
    }
  }
  const long n1 = A.g.n[1], n2 = A.g.n[2];
  const long x0 = (long)T0 * TS + q0, x1 = (long)T1 * TS + q1, x2 = (long)T2 * TS + c;
  if (A.c2r) {  // real grid of a c2r plan
    R* dr = reinterpret_cast<R*>(A.grid) + (x0 * n1 + x1) * n2 + x2;
    dr[0] = (R)sr[0];
    dr[1] = (R)sr[1];
    return;
  }
  Cplx<R>* dst = A.grid + (x0 * n1 + x1) * n2 + x2;
  if constexpr (sizeof(R) == 4) {
    *reinterpret_cast<float4*>(dst) = make_float4((float)sr[0], (float)si[0], (float)sr[1], (float)si[1]);
  } else {
#pragma unroll
    for (int k = 0; k < 2; k++) dst[k] = cmk<R>((R)sr[k], (R)si[k]);
  }
}

// nodes whose stencil start left the tile they were binned to (bins older
// than the coordinates), listed by the column kernel: added to the folded grid
// through global atomics
template <class R, int MT, int MODE>
__global__ __launch_bounds__(256) void scatter_stale_k(BArgs<R> A) {
  using S = ColsShape<R, MT, MODE>;
  constexpr int CM = S::CM;
  const unsigned count = *A.stale_count;
  for (unsigned k = blockIdx.x * 256 + threadIdx.x; k < count; k += gridDim.x * 256) {
    const int p = A.stale_list[k];
    const long j = A.perm[p];
    NodeWin<R, CM> w;
    node_windows<R, MT, CM, S::WG>(A, p, j, MT, w);
    scatter_node_global<R, CM, MODE, true>(A, j, w, CM);
  }
}

__global__ void zero_count_k(unsigned* c) {
  if (threadIdx.x == 0) *c = 0u;
}

inline unsigned grid1d(long n, int t) { return (unsigned)((n + t - 1) / t); }

template <class R, int MT, int MODE>
void launch_m(const BArgs<R>& A0, hipStream_t s) {
  using S = ColsShape<R, MT, MODE>;
  // box slots where the matrix-core scatter writes the slots and fold8b_k reads them
  BArgs<R> A = A0;
  A.box_slots = 0;
  if constexpr (sizeof(R) == 4 && S::TS == 8 && S::P <= 16 && S::P % 4 == 0)
    A.box_slots = (A.box_tile && A.box_ord && !A.no_mfma && A.ts[2] == 8 && A.ntile[2] % 2 == 0) ? 1 : 0;
  hipLaunchKernelGGL(zero_count_k, dim3(1), dim3(64), 0, s, A.stale_count);
  bool done = false;
  if constexpr (sizeof(R) == 4 && S::P <= 16) {
    if (!A.no_mfma) {
      if (A.c2r)
        hipLaunchKernelGGL((scatter_mfma_k<MT, MODE, true>), dim3((unsigned)A.n_torder), dim3(256), 0, s, A);
      else
        hipLaunchKernelGGL((scatter_mfma_k<MT, MODE, false>), dim3((unsigned)A.n_torder), dim3(256), 0, s, A);
      done = true;
    }
  }
  if (!done) hipLaunchKernelGGL((scatter_cols_k<R, MT, MODE>), dim3((unsigned)A.n_torder), dim3(S::NT), 0, s, A);
  if (A.mid) (void)hipEventRecord(A.mid, s);
  const long ntiles = (long)A.ntile[0] * A.ntile[1] * A.ntile[2];
  if constexpr (S::TS == 8) {
    bool brick = false;
    if constexpr (S::P <= 16 && S::P % 4 == 0) {
      if (A.ts[2] == 8 && A.ntile[2] % 2 == 0) {
        hipLaunchKernelGGL((fold8b_k<R, S::P>), dim3((unsigned)(ntiles / 2)), dim3(256), 0, s, A);
        brick = true;
      }
    }
    if (!brick) hipLaunchKernelGGL((fold8_k<R, S::P>), dim3((unsigned)ntiles), dim3(128), 0, s, A);
  } else if constexpr (S::P <= 16) {
    if (A.ts[2] == 4 && A.ntile[2] % 4 == 0)
      hipLaunchKernelGGL((fold4_k<R, S::P>), dim3((unsigned)(ntiles / 4)), dim3(128), 0, s, A);
    else
      hipLaunchKernelGGL((fold_k<R, S::TS, S::P>), dim3((unsigned)ntiles), dim3(S::TS * S::TS * S::TS), 0, s, A);
  } else {
    hipLaunchKernelGGL((fold_k<R, S::TS, S::P>), dim3((unsigned)ntiles), dim3(S::TS * S::TS * S::TS), 0, s, A);
  }
  if (A.mid2) (void)hipEventRecord(A.mid2, s);
  hipLaunchKernelGGL((scatter_stale_k<R, MT, MODE>), dim3(256), dim3(256), 0, s, A);
}

}  // namespace

template <class R, int MODE>
int cols_launch(const BArgs<R>& A, hipStream_t s) {
  if (A.M == 0 || A.n_torder <= 0 || !A.stale_count || !A.stale_list) return kUnsupported;
  // FP64 gradient modes would need more than 256 VGPRs per lane here: they keep the LDS-atomic path
  if constexpr (sizeof(R) == 8 && MODE != B_F) {
    return kUnsupported;
  } else {
    switch (A.g.m) {
      case 2: launch_m<R, 2, MODE>(A, s); break;
      case 3: launch_m<R, 3, MODE>(A, s); break;
      case 4: launch_m<R, 4, MODE>(A, s); break;
      case 5: launch_m<R, 5, MODE>(A, s); break;
      case 6: launch_m<R, 6, MODE>(A, s); break;
      case 7:
        if constexpr (sizeof(R) != 4) return kUnsupported;
        else launch_m<R, 7, MODE>(A, s);
        break;
      case 8:
        if constexpr (sizeof(R) != 4) return kUnsupported;
        else launch_m<R, 8, MODE>(A, s);
        break;
      default: return kUnsupported;
    }
    return hipGetLastError() == hipSuccess ? kOk : kHipError;
  }
}

template int cols_launch<PNFFT_B_REAL, PNFFT_B_MODE>(const BArgs<PNFFT_B_REAL>&, hipStream_t);

}  // namespace pnfft_amd
