// b_device.hpp -- device helpers shared by the B kernels (b_kernels.hip) and
// the column scatter (scatter_cols.hip): per-node window evaluation, node
// value loads, the separable gather sums and the per-node global scatter.
#pragma once

#include "b_common.hpp"
#include "device_common.hpp"
#include "nfft_b.hpp"
#include "window_math.hpp"

namespace pnfft_amd {
namespace bdev {

constexpr int kB = 256;  // threads per block, direct kernels

// Kernel bodies with compile-time stencil loops (window arrays in registers)
// are used for FP32 at every m and for FP64 up to m = 6; FP64 at m = 7, 8 keeps
// the runtime-width loops (the static bodies there need more than 256 VGPRs).
template <class R, int MT>
constexpr bool static_body() {
  return MT > 0 && (sizeof(R) == 4 || MT <= 6);
}

template <class R, int CM>
struct NodeWin {
  int start[3];  // natural index of stencil point 0 per axis
  R px[CM], py[CM], pz[CM];
  R dx[CM], dy[CM], dz[CM];
};

// Multiply the window values of each axis by the plan's power-of-two scale
// (GridGeom::ws; exact), c <= CM entries.
template <class R, int CM, bool WG>
__device__ inline void scale_windows(const GridGeom& g, int c, NodeWin<R, CM>& w) {
  const R s0 = (R)g.ws[0], s1 = (R)g.ws[1], s2 = (R)g.ws[2];
#pragma unroll
  for (int k = 0; k < CM; k++) {
    if (k < c) {
      w.px[k] *= s0;
      w.py[k] *= s1;
      w.pz[k] *= s2;
      if (WG) {
        w.dx[k] *= s0;
        w.dy[k] *= s1;
        w.dz[k] *= s2;
      }
    }
  }
}

// Window values of one node (node id j, position p in the sorted order).
// MT > 0: the width CM = 2 MT + 1 is a compile-time constant and every array
// index is static, so NodeWin lives in registers; the values come from the
// precomputed tables (pnfft_precompute_psi) when they exist (with the
// derivatives, if WG) and the node's coordinates still equal the ones they were
// computed from.  MT == 0: the runtime-width loops of the FP64 m = 7, 8
// kernels (width m, arrays of CM entries), always evaluated.
template <class R, int MT, int CM, bool WG>
__device__ inline void node_windows(const BArgs<R>& A, long p, long j, int m, NodeWin<R, CM>& w) {
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const R x0 = A.x[3 * j + 0], x1 = A.x[3 * j + 1], x2 = A.x[3 * j + 2];
  const NodeAxis<R> a0 = node_axis<R>(x0, n0, m, A.g.shift);
  const NodeAxis<R> a1 = node_axis<R>(x1, n1, m, A.g.shift);
  const NodeAxis<R> a2 = node_axis<R>(x2, n2, m, A.g.shift);
  w.start[0] = a0.start;
  w.start[1] = a1.start;
  w.start[2] = a2.start;
  if constexpr (MT > 0) {
    // table values are loaded before the coordinate check, so that both round
    // trips overlap; they are kept when x is unchanged since the tables were made
    bool have = false;
    if (A.pre_psi && p >= 0) {
      const long M = A.M;
#pragma unroll
      for (int s = 0; s < CM; s++) {
        w.px[s] = A.pre_psi[(long)s * M + p];
        w.py[s] = A.pre_psi[(long)(CM + s) * M + p];
        w.pz[s] = A.pre_psi[(long)(2 * CM + s) * M + p];
      }
      if (WG && A.pre_dpsi) {
#pragma unroll
        for (int s = 0; s < CM; s++) {
          w.dx[s] = A.pre_dpsi[(long)s * M + p];
          w.dy[s] = A.pre_dpsi[(long)(CM + s) * M + p];
          w.dz[s] = A.pre_dpsi[(long)(2 * CM + s) * M + p];
        }
      }
      const bool same = x0 == A.pre_x[p] && x1 == A.pre_x[M + p] && x2 == A.pre_x[2 * M + p];
      have = same && (!WG || A.pre_dpsi != nullptr);
    }
    if (!have) {
      window_axis_s<R, CM, WG>(A.g.win, a0.frac, (R)n0, (R)A.g.b[0], w.px, w.dx);
      window_axis_s<R, CM, WG>(A.g.win, a1.frac, (R)n1, (R)A.g.b[1], w.py, w.dy);
      window_axis_s<R, CM, WG>(A.g.win, a2.frac, (R)n2, (R)A.g.b[2], w.pz, w.dz);
      scale_windows<R, CM, WG>(A.g, CM, w);
    }
  } else {
    window_axis<R>(A.g.win, m, a0.frac, (R)n0, (R)A.g.b[0], w.px, WG ? w.dx : nullptr);
    window_axis<R>(A.g.win, m, a1.frac, (R)n1, (R)A.g.b[1], w.py, WG ? w.dy : nullptr);
    window_axis<R>(A.g.win, m, a2.frac, (R)n2, (R)A.g.b[2], w.pz, WG ? w.dz : nullptr);
    scale_windows<R, CM, WG>(A.g, 2 * m + 1, w);
  }
}

// Drop the x-window entry of plane a: after the shift, px[0] / dx[0] belong to
// plane a + 1.  Lets the plane loop stay rolled while every index is static.
template <class R, int CM, bool WG>
__device__ inline void shift_x(NodeWin<R, CM>& w) {
#pragma unroll
  for (int k = 0; k + 1 < CM; k++) {
    w.px[k] = w.px[k + 1];
    if (WG) w.dx[k] = w.dx[k + 1];
  }
}

// Separable gather sum f = sum_a px sum_b py sum_c pz G (and the three
// gradient sums with one factor replaced by its derivative), static width CM.
// ROW(a, b) returns a pointer p with p[IDX(c)] the grid value at stencil (a,b,c).
// a grid cell read through a row pointer: complex grids, or the real grid of a
// c2r plan (imaginary part zero)
template <class R>
__device__ __forceinline__ Cplx<R> grid_cell(const Cplx<R>* row, int i) { return row[i]; }
template <class R>
__device__ __forceinline__ Cplx<R> grid_cell(const R* row, int i) { return cmk<R>(row[i], R(0)); }

template <class R, int CM, int MODE, class RowFn, class IdxFn>
__device__ inline void gather_sum(NodeWin<R, CM> w, RowFn row_of, IdxFn idx, Cplx<R>* out) {
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD || MODE == B_VJP);
  Cplx<R> f = cmk<R>(0, 0), g0 = cmk<R>(0, 0), g1 = cmk<R>(0, 0), g2 = cmk<R>(0, 0);
#pragma unroll 1
  for (int a = 0; a < CM; a++) {
    Cplx<R> ha = cmk<R>(0, 0), hay = cmk<R>(0, 0), haz = cmk<R>(0, 0);
#pragma unroll
    for (int b = 0; b < CM; b++) {
      const auto row = row_of(a, b);
      Cplx<R> hb = cmk<R>(0, 0), hbz = cmk<R>(0, 0);
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const Cplx<R> gv = grid_cell<R>(row, idx(s));
        hb.x += w.pz[s] * gv.x;
        hb.y += w.pz[s] * gv.y;
        if (WG) {
          hbz.x += w.dz[s] * gv.x;
          hbz.y += w.dz[s] * gv.y;
        }
      }
      ha.x += w.py[b] * hb.x;
      ha.y += w.py[b] * hb.y;
      if (WG) {
        hay.x += w.dy[b] * hb.x;
        hay.y += w.dy[b] * hb.y;
        haz.x += w.py[b] * hbz.x;
        haz.y += w.py[b] * hbz.y;
      }
    }
    f.x += w.px[0] * ha.x;
    f.y += w.px[0] * ha.y;
    if (WG) {
      g0.x += w.dx[0] * ha.x;
      g0.y += w.dx[0] * ha.y;
      g1.x += w.px[0] * hay.x;
      g1.y += w.px[0] * hay.y;
      g2.x += w.px[0] * haz.x;
      g2.y += w.px[0] * haz.y;
    }
    shift_x<R, CM, WG>(w);
  }
  out[0] = f;
  out[1] = g0;
  out[2] = g1;
  out[3] = g2;
}

// The same sum for a runtime width c <= CM (generic m); same operation order.
template <class R, int CM, int MODE, class RowFn, class IdxFn>
__device__ inline void gather_sum_rt(const NodeWin<R, CM>& w, int c, RowFn row_of, IdxFn idx, Cplx<R>* out) {
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD || MODE == B_VJP);
  Cplx<R> f = cmk<R>(0, 0), g0 = cmk<R>(0, 0), g1 = cmk<R>(0, 0), g2 = cmk<R>(0, 0);
  for (int a = 0; a < c; a++) {
    Cplx<R> ha = cmk<R>(0, 0), hay = cmk<R>(0, 0), haz = cmk<R>(0, 0);
    for (int b = 0; b < c; b++) {
      const auto row = row_of(a, b);
      Cplx<R> hb = cmk<R>(0, 0), hbz = cmk<R>(0, 0);
      for (int s = 0; s < c; s++) {
        const Cplx<R> gv = grid_cell<R>(row, idx(s));
        hb.x += w.pz[s] * gv.x;
        hb.y += w.pz[s] * gv.y;
        if (WG) {
          hbz.x += w.dz[s] * gv.x;
          hbz.y += w.dz[s] * gv.y;
        }
      }
      ha.x += w.py[b] * hb.x;
      ha.y += w.py[b] * hb.y;
      if (WG) {
        hay.x += w.dy[b] * hb.x;
        hay.y += w.dy[b] * hb.y;
        haz.x += w.py[b] * hbz.x;
        haz.y += w.py[b] * hbz.y;
      }
    }
    f.x += w.px[a] * ha.x;
    f.y += w.px[a] * ha.y;
    if (WG) {
      g0.x += w.dx[a] * ha.x;
      g0.y += w.dx[a] * ha.y;
      g1.x += w.px[a] * hay.x;
      g1.y += w.px[a] * hay.y;
      g2.x += w.px[a] * haz.x;
      g2.y += w.px[a] * haz.y;
    }
  }
  out[0] = f;
  out[1] = g0;
  out[2] = g1;
  out[3] = g2;
}

// node values: complex arrays, or real arrays of a c2r plan (imaginary part 0)
template <class R>
__device__ inline Cplx<R> load_f(const BArgs<R>& A, long j) {
  return A.c2r ? cmk<R>(reinterpret_cast<const R*>(A.f)[j], R(0)) : A.f[j];
}
template <class R>
__device__ inline Cplx<R> load_g(const BArgs<R>& A, long j, int t) {
  return A.c2r ? cmk<R>(reinterpret_cast<const R*>(A.grad)[3 * j + t], R(0)) : A.grad[3 * j + t];
}

template <class R, int MODE>
__device__ inline void gather_store(const BArgs<R>& A, long j, Cplx<R>* v) {
  if (A.c2r) {  // real plan: f = Re(...), grad = Re(...) (kernel/assign.c:715-739, :821-851)
    R* fr = reinterpret_cast<R*>(A.f);
    R* gr = reinterpret_cast<R*>(A.grad);
    if (MODE == B_F || MODE == B_FGRAD) fr[j] = (A.accumulate ? fr[j] : R(0)) + v[0].x;
    if (MODE == B_GRAD || MODE == B_FGRAD)
      for (int t = 0; t < 3; t++) gr[3 * j + t] = (A.accumulate ? gr[3 * j + t] : R(0)) + v[1 + t].x;
    if (MODE == B_VJP) {
      const R wv = reinterpret_cast<const R*>(A.w)[j];
      for (int t = 0; t < 3; t++) A.gx[3 * j + t] = wv * v[1 + t].x;
    }
    return;
  }
  if (MODE == B_F || MODE == B_FGRAD) {
    Cplx<R> o = A.accumulate ? A.f[j] : cmk<R>(0, 0);
    Cplx<R> f = v[0];
    if (A.real_only) f.y = 0;  // assign_f_r2r reads the real grid part only (kernel/assign.c:715-739)
    A.f[j] = cadd(o, f);
  }
  if (MODE == B_GRAD || MODE == B_FGRAD) {
    Cplx<R>* gp = A.grad + 3 * j;
    if (A.real_only) {  // assign_grad_f_r2r writes the real parts only (kernel/assign.c:821-851)
      v[1].y = 0;
      v[2].y = 0;
      v[3].y = 0;
    }
    if (A.accumulate) {
      v[1] = cadd(v[1], gp[0]);
      v[2] = cadd(v[2], gp[1]);
      v[3] = cadd(v[3], gp[2]);
    }
    gp[0] = v[1];
    gp[1] = v[2];
    gp[2] = v[3];
  }
  if (MODE == B_VJP) {
    const Cplx<R> wv = A.w[j];
    A.gx[3 * j + 0] = wv.x * v[1].x + wv.y * v[1].y;
    A.gx[3 * j + 1] = wv.x * v[2].x + wv.y * v[2].y;
    A.gx[3 * j + 2] = wv.x * v[3].x + wv.y * v[3].y;
  }
}

// gather one node straight from the grid (periodic wrap on every axis).
// ST: the stencil width is the compile-time CM (else the runtime c <= CM).
template <class R, int CM, int MODE, bool ST>
__device__ inline void gather_node_global(const BArgs<R>& A, long j, const NodeWin<R, CM>& w, int c) {
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  int i2s[CM];
#pragma unroll
  for (int s = 0; s < CM; s++) {
    const int v = w.start[2] + s;
    i2s[s] = v >= n2 ? v - n2 : v;
  }
  auto row_off = [&](int a, int b) {
    int i0 = w.start[0] + a;
    if (i0 >= n0) i0 -= n0;
    int i1 = w.start[1] + b;
    if (i1 >= n1) i1 -= n1;
    return ((long)i0 * n1 + i1) * n2;
  };
  auto idx = [&](int s) { return i2s[s]; };
  Cplx<R> v[4];
  if (A.c2r) {  // real grid
    const R* grid = reinterpret_cast<const R*>(A.grid);
    auto row_of = [&](int a, int b) { return grid + row_off(a, b); };
    if constexpr (ST)
      gather_sum<R, CM, MODE>(w, row_of, idx, v);
    else
      gather_sum_rt<R, CM, MODE>(w, c, row_of, idx, v);
  } else {
    const Cplx<R>* grid = A.grid;
    auto row_of = [&](int a, int b) { return grid + row_off(a, b); };
    if constexpr (ST)
      gather_sum<R, CM, MODE>(w, row_of, idx, v);
    else
      gather_sum_rt<R, CM, MODE>(w, c, row_of, idx, v);
  }
  gather_store<R, MODE>(A, j, v);
}

// value added at one stencil point: f psi(a) psi(b) psi(c) and/or
// grad . (psi'(a) psi(b) psi(c), psi(a) psi'(b) psi(c), psi(a) psi(b) psi'(c))
template <class R, bool WF, bool WG>
__device__ inline Cplx<R> scatter_value(R pxy, R pdxy, R pxdy, R pz, R dz, Cplx<R> fv, Cplx<R> gv0, Cplx<R> gv1,
                                        Cplx<R> gv2) {
  R vr = 0, vi = 0;
  if (WF) {
    const R q = pxy * pz;
    vr += q * fv.x;
    vi += q * fv.y;
  }
  if (WG) {
    const R q0 = pdxy * pz, q1 = pxdy * pz, q2 = pxy * dz;
    vr += q0 * gv0.x + q1 * gv1.x + q2 * gv2.x;
    vi += q0 * gv0.y + q1 * gv1.y + q2 * gv2.y;
  }
  return cmk<R>(vr, vi);
}

// adds v at cell row + i of the grid (global atomics): both parts, or the real
// part of a c2r plan's real grid (real node values leave no imaginary part)
template <class R>
__device__ __forceinline__ void add_cell(const BArgs<R>& A, long row, int i, Cplx<R> v) {
  if (A.c2r) {
    atomicAdd(reinterpret_cast<R*>(A.grid) + row + i, v.x);
  } else {
    atomicAdd(&A.grid[row + i].x, v.x);
    atomicAdd(&A.grid[row + i].y, v.y);
  }
}

template <class R, int CM, int MODE, bool ST>
__device__ inline void scatter_node_global(const BArgs<R>& A, long j, NodeWin<R, CM> w, int c) {
  constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  Cplx<R> fv = cmk<R>(0, 0), gv0 = cmk<R>(0, 0), gv1 = cmk<R>(0, 0), gv2 = cmk<R>(0, 0);
  if (WF) fv = load_f(A, j);
  if (WG) {
    gv0 = load_g(A, j, 0);
    gv1 = load_g(A, j, 1);
    gv2 = load_g(A, j, 2);
  }
  if constexpr (ST) {
    int i2s[CM];
#pragma unroll
    for (int s = 0; s < CM; s++) {
      const int v = w.start[2] + s;
      i2s[s] = v >= n2 ? v - n2 : v;
    }
#pragma unroll 1
    for (int a = 0; a < CM; a++) {
      int i0 = w.start[0] + a;
      if (i0 >= n0) i0 -= n0;
#pragma unroll
      for (int b = 0; b < CM; b++) {
        int i1 = w.start[1] + b;
        if (i1 >= n1) i1 -= n1;
        const long row = ((long)i0 * n1 + i1) * n2;
        const R pxy = w.px[0] * w.py[b];
        const R pdxy = WG ? w.dx[0] * w.py[b] : R(0);
        const R pxdy = WG ? w.px[0] * w.dy[b] : R(0);
#pragma unroll
        for (int s = 0; s < CM; s++) {
          const Cplx<R> v = scatter_value<R, WF, WG>(pxy, pdxy, pxdy, w.pz[s], WG ? w.dz[s] : R(0), fv, gv0, gv1, gv2);
          add_cell<R>(A, row, i2s[s], v);
        }
      }
      shift_x<R, CM, WG>(w);
    }
  } else {
    for (int a = 0; a < c; a++) {
      int i0 = w.start[0] + a;
      if (i0 >= n0) i0 -= n0;
      for (int b = 0; b < c; b++) {
        int i1 = w.start[1] + b;
        if (i1 >= n1) i1 -= n1;
        const long row = ((long)i0 * n1 + i1) * n2;
        const R pxy = w.px[a] * w.py[b];
        const R pdxy = WG ? w.dx[a] * w.py[b] : R(0);
        const R pxdy = WG ? w.px[a] * w.dy[b] : R(0);
        for (int s = 0; s < c; s++) {
          int i2 = w.start[2] + s;
          if (i2 >= n2) i2 -= n2;
          const Cplx<R> v = scatter_value<R, WF, WG>(pxy, pdxy, pxdy, w.pz[s], WG ? w.dz[s] : R(0), fv, gv0, gv1, gv2);
          add_cell<R>(A, row, i2, v);
        }
      }
    }
  }
}

__device__ inline int wrapi(int v, int n) { return v >= n ? v - n : v; }

// local offset of the node's stencil start inside the tile, or -1 if outside
__device__ inline int local_off(int start, int o, int ts, int n) {
  int l = start - o;
  if (l < 0) l += n;
  return (l < ts) ? l : -1;
}

}  // namespace bdev
}  // namespace pnfft_amd
