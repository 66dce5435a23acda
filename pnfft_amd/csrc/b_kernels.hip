// b_kernels.hip -- gather (trafo B) and scatter (adj B^T) kernels.
//
// This file is compiled once per (precision, kind, mode); the build passes
// PNFFT_B_REAL (float|double), PNFFT_B_KIND (0 gather, 1 scatter) and
// PNFFT_B_MODE (BMode), and the file instantiates the launcher
// b_launch<PNFFT_B_REAL, PNFFT_B_KIND, PNFFT_B_MODE> for them.
//
// Reference: the node loops of kernel/ndft-parallel.c:2703-3009 and the
// stencil kernels of kernel/assign.c:478-1066 (cutoff 2m+1, :910).
//
// Two paths:
//  * tiled: one block per work item (tile, chunk of <= kTileNodes sorted
//    nodes).  The tile's padded subgrid ((TS + 2m) cells per axis, starting at
//    the tile origin because tiles are keyed by the stencil start) is staged in
//    LDS; the gather reads every stencil point from LDS, the scatter
//    accumulates into LDS in FP64 (ds_add_f64) and adds the nonzero cells of
//    the subgrid to the grid once (global float atomics), so overlapping halos
//    of neighbouring tiles meet in L2 instead of per stencil point.
//  * direct: one thread per node, stencil read from / added to the grid in
//    HBM; used for nodes of sparse tiles and for any node that does not lie in
//    the tile it was binned to (bins older than the node coordinates).
#include "b_device.hpp"

namespace pnfft_amd {

namespace {

using namespace bdev;

// ---------------------------------------------------------------------------
// direct kernels: positions [0, M) of the sorted order; with tiles present,
// only the nodes of sparse tiles (tile_dense[tile] == 0) are handled here.
// ---------------------------------------------------------------------------
template <class R>
__device__ inline bool skip_dense(const BArgs<R>& A, long p) {
  if (!A.tile_dense) return false;
  const unsigned tile = A.keys[p] / (unsigned)(A.ts[0] * A.ts[1] * A.ts[2]);
  return A.tile_dense[tile] != 0;
}

template <class R, int MT, int MODE>
__global__ __launch_bounds__(kB) void gather_direct_k(BArgs<R> A) {
  const long p = (long)blockIdx.x * kB + threadIdx.x;
  if (p >= A.M) return;
  if (skip_dense(A, p)) return;
  const long j = A.perm ? (long)A.perm[p] : p;
  constexpr int m = MT;
  constexpr int CM = 2 * MT + 1;
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD || MODE == B_VJP);
  NodeWin<R, CM> w;
  constexpr bool ST = static_body<R, MT>();
  node_windows<R, ST ? MT : 0, CM, WG>(A, A.perm ? p : -1, j, m, w);
  gather_node_global<R, CM, MODE, ST>(A, j, w, 2 * m + 1);
}

template <class R, int MT, int MODE>
__global__ __launch_bounds__(kB) void scatter_direct_k(BArgs<R> A) {
  const long p = (long)blockIdx.x * kB + threadIdx.x;
  if (p >= A.M) return;
  if (skip_dense(A, p)) return;
  const long j = A.perm ? (long)A.perm[p] : p;
  constexpr int m = MT;
  constexpr int CM = 2 * MT + 1;
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  NodeWin<R, CM> w;
  constexpr bool ST = static_body<R, MT>();
  node_windows<R, ST ? MT : 0, CM, WG>(A, A.perm ? p : -1, j, m, w);
  scatter_node_global<R, CM, MODE, ST>(A, j, w, 2 * m + 1);
}

// ---------------------------------------------------------------------------
// tiled kernels
// ---------------------------------------------------------------------------
struct TileItem {
  int tile;
  int o[3];      // tile origin (natural grid index)
  int P[3];      // padded extent TS + 2m
  int beg, end;  // node range in the sorted order
};

template <class R>
__device__ inline TileItem tile_item(const BArgs<R>& A, int m) {
  TileItem it;
  const int t = A.items[2 * blockIdx.x];
  const int chunk = A.items[2 * blockIdx.x + 1];
  it.tile = t;
  const int t2 = t % A.ntile[2];
  const int t1 = (t / A.ntile[2]) % A.ntile[1];
  const int t0 = t / (A.ntile[2] * A.ntile[1]);
  it.o[0] = t0 * A.ts[0];
  it.o[1] = t1 * A.ts[1];
  it.o[2] = t2 * A.ts[2];
  for (int d = 0; d < 3; d++) it.P[d] = A.ts[d] + 2 * m;
  const int s = A.tile_start[t], e = A.tile_start[t + 1];
  it.beg = s + chunk * kTileNodes;
  it.end = min(e, it.beg + kTileNodes);
  return it;
}

// LDS subgrid layout: cell (a, b, c) at a*PP + b*RP + c with an odd row pitch
// RP = P2 + 1 and plane pitch PP = P1*RP + 1, so that rows and planes of the
// subgrid start in different banks.
struct SubgridPitch {
  int RP, PP;
};
__device__ inline SubgridPitch subgrid_pitch(const TileItem& it) {
  SubgridPitch s;
  s.RP = it.P[2] + 1;
  s.PP = it.P[1] * s.RP + 1;
  return s;
}

// Copy of a tile's padded subgrid (P^3 cells from the tile origin, periodic
// wrap) into LDS, for the standard layout (tile size TS on every axis, so P =
// TS + 2m is a compile-time constant).  Each thread moves pairs of cells along
// axis 2: the origin is a multiple of TS and n2 is a power of two, so a pair
// starting at an even offset never straddles the wrap and is one aligned load.
template <class R, int P>
__device__ inline void stage_subgrid_static(const BArgs<R>& A, const TileItem& it, Cplx<R>* tile) {
  static_assert(P % 2 == 0, "P = TS + 2m is even");
  constexpr int H = P / 2, RP = P + 1, PP = P * RP + 1, NPAIR = P * P * H;
  constexpr int K = (NPAIR + kTileThreads - 1) / kTileThreads;
  using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;  // two cells
  using V2 = typename std::conditional<sizeof(R) == 4, float2, double2>::type;  // two cells of a real grid
  const unsigned n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const V* src = reinterpret_cast<const V*>(A.grid);
  const V2* src_r = reinterpret_cast<const V2*>(A.grid);
  // all loads first, then the LDS stores: the K round trips overlap
  V v[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int q = threadIdx.x + k * kTileThreads;
    if (q < NPAIR) {
      const int hz = q % H, cy = (q / H) % P, cx = q / (H * P);
      const unsigned g0 = (unsigned)(it.o[0] + cx) & (n0 - 1), g1 = (unsigned)(it.o[1] + cy) & (n1 - 1);
      const unsigned g2 = (unsigned)(it.o[2] + 2 * hz) & (n2 - 1);
      if (A.c2r) {  // real grid (c2r plans): imaginary parts zero
        const V2 r = src_r[((g0 * n1 + g1) * n2 + g2) >> 1];
        v[k].x = r.x;
        v[k].y = R(0);
        v[k].z = r.y;
        v[k].w = R(0);
      } else {
        v[k] = src[((g0 * n1 + g1) * n2 + g2) >> 1];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int q = threadIdx.x + k * kTileThreads;
    if (q < NPAIR) {
      const int hz = q % H, cy = (q / H) % P, cx = q / (H * P);
      Cplx<R>* d = tile + cx * PP + cy * RP + 2 * hz;
      d[0] = cmk<R>(v[k].x, v[k].y);
      d[1] = cmk<R>(v[k].z, v[k].w);
    }
  }
}

// Waves per SIMD the FP32 gather kernels are compiled for (register budget
// 512 / w): 4 for f, 3 for the gradient modes; FP64 keeps the compiler's
// choice (a budget would make it spill).
template <class R, int MODE>
constexpr int gather_waves() {
  return sizeof(R) == 4 ? (MODE == B_F ? 4 : 3) : 1;
}

template <class R, int MT, int MODE>
__global__ __launch_bounds__(kTileThreads, (gather_waves<R, MODE>())) void gather_tiled_k(BArgs<R> A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<R>* tile = reinterpret_cast<Cplx<R>*>(smem_raw);
  constexpr int m = MT;
  constexpr int CM = 2 * MT + 1;
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD || MODE == B_VJP);
  constexpr bool ST = static_body<R, MT>();
  const TileItem it = tile_item(A, m);
  const SubgridPitch sp = subgrid_pitch(it);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const int P1 = it.P[1], P2 = it.P[2];
  const int vol = it.P[0] * P1 * P2;
  constexpr int TSL = std_tile(MT, sizeof(R) == 8);  // tile size of the standard layout
  if (ST && A.ts[0] == TSL && A.ts[1] == TSL && A.ts[2] == TSL) {
    stage_subgrid_static<R, TSL + 2 * MT>(A, it, tile);
  } else {
    for (int idx = threadIdx.x; idx < vol; idx += kTileThreads) {
      const int cz = idx % P2, cy = (idx / P2) % P1, cx = idx / (P1 * P2);
      const int g0 = wrapi(it.o[0] + cx, n0), g1 = wrapi(it.o[1] + cy, n1), g2 = wrapi(it.o[2] + cz, n2);
      const long gi = ((long)g0 * n1 + g1) * n2 + g2;
      tile[cx * sp.PP + cy * sp.RP + cz] = A.c2r ? cmk<R>(reinterpret_cast<const R*>(A.grid)[gi], R(0)) : A.grid[gi];
    }
  }
  __syncthreads();
  for (int p = it.beg + threadIdx.x; p < it.end; p += kTileThreads) {
    const long j = A.perm[p];
    NodeWin<R, CM> w;
    node_windows<R, ST ? MT : 0, CM, WG>(A, A.perm ? p : -1, j, m, w);
    const int l0 = local_off(w.start[0], it.o[0], A.ts[0], n0);
    const int l1 = local_off(w.start[1], it.o[1], A.ts[1], n1);
    const int l2 = local_off(w.start[2], it.o[2], A.ts[2], n2);
    if (l0 < 0 || l1 < 0 || l2 < 0) {
      gather_node_global<R, CM, MODE, ST>(A, j, w, CM);
      continue;
    }
    const Cplx<R>* base = tile + l0 * sp.PP + l1 * sp.RP + l2;
    auto row_of = [&](int a, int b) { return base + a * sp.PP + b * sp.RP; };
    auto idx = [](int s) { return s; };
    Cplx<R> v[4];
    if constexpr (ST)
      gather_sum<R, CM, MODE>(w, row_of, idx, v);
    else
      gather_sum_rt<R, CM, MODE>(w, CM, row_of, idx, v);
    gather_store<R, MODE>(A, j, v);
  }
}

template <class R, int MT, int MODE>
__global__ __launch_bounds__(kTileThreads) void scatter_tiled_k(BArgs<R> A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* acc = reinterpret_cast<double*>(smem_raw);  // interleaved (re, im) FP64 accumulators
  constexpr int m = MT;
  constexpr int CM = 2 * MT + 1;
  constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  constexpr bool ST = static_body<R, MT>();
  const TileItem it = tile_item(A, m);
  const SubgridPitch sp = subgrid_pitch(it);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const int P1 = it.P[1], P2 = it.P[2];
  const int vol = it.P[0] * P1 * P2;
  for (int idx = threadIdx.x; idx < 2 * it.P[0] * sp.PP; idx += kTileThreads) acc[idx] = 0.0;
  __syncthreads();
  // Sorted nodes that share a cell are consecutive; dealing them out across the
  // waves (lane-major) keeps the lanes of one wave on different cells, so one
  // ds_add instruction rarely hits the same LDS address twice.
  constexpr int kWaves = kTileThreads / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = lane * kWaves + wave; it.beg + q < it.end; q += kTileThreads) {
    const int p = it.beg + q;
    const long j = A.perm[p];
    NodeWin<R, CM> w;
    node_windows<R, ST ? MT : 0, CM, WG>(A, A.perm ? p : -1, j, m, w);
    const int l0 = local_off(w.start[0], it.o[0], A.ts[0], n0);
    const int l1 = local_off(w.start[1], it.o[1], A.ts[1], n1);
    const int l2 = local_off(w.start[2], it.o[2], A.ts[2], n2);
    if (l0 < 0 || l1 < 0 || l2 < 0) {
      scatter_node_global<R, CM, MODE, ST>(A, j, w, CM);
      continue;
    }
    Cplx<R> fv = cmk<R>(0, 0), gv0 = cmk<R>(0, 0), gv1 = cmk<R>(0, 0), gv2 = cmk<R>(0, 0);
    if (WF) fv = load_f(A, j);
    if (WG) {
      gv0 = load_g(A, j, 0);
      gv1 = load_g(A, j, 1);
      gv2 = load_g(A, j, 2);
    }
    double* plane = acc + 2 * (l0 * sp.PP + l1 * sp.RP + l2);
    if constexpr (ST) {
#pragma unroll 1
      for (int a = 0; a < CM; a++, plane += 2 * sp.PP) {
#pragma unroll
        for (int b = 0; b < CM; b++) {
          double* row = plane + 2 * b * sp.RP;
          const R pxy = w.px[0] * w.py[b];
          const R pdxy = WG ? w.dx[0] * w.py[b] : R(0);
          const R pxdy = WG ? w.px[0] * w.dy[b] : R(0);
#pragma unroll
          for (int s = 0; s < CM; s++) {
            const Cplx<R> v =
                scatter_value<R, WF, WG>(pxy, pdxy, pxdy, w.pz[s], WG ? w.dz[s] : R(0), fv, gv0, gv1, gv2);
            atomicAdd(&row[2 * s], (double)v.x);
            atomicAdd(&row[2 * s + 1], (double)v.y);
          }
        }
        shift_x<R, CM, WG>(w);
      }
    } else {
      for (int a = 0; a < CM; a++) {
        for (int b = 0; b < CM; b++) {
          double* row = plane + 2 * (a * sp.PP + b * sp.RP);
          const R pxy = w.px[a] * w.py[b];
          const R pdxy = WG ? w.dx[a] * w.py[b] : R(0);
          const R pxdy = WG ? w.px[a] * w.dy[b] : R(0);
#pragma unroll
          for (int s = 0; s < CM; s++) {
            const Cplx<R> v =
                scatter_value<R, WF, WG>(pxy, pdxy, pxdy, w.pz[s], WG ? w.dz[s] : R(0), fv, gv0, gv1, gv2);
            atomicAdd(&row[2 * s], (double)v.x);
            atomicAdd(&row[2 * s + 1], (double)v.y);
          }
        }
      }
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < vol; idx += kTileThreads) {
    const int cz = idx % P2, cy = (idx / P2) % P1, cx = idx / (P1 * P2);
    const int li = cx * sp.PP + cy * sp.RP + cz;
    const double vr = acc[2 * li], vi = acc[2 * li + 1];
    if (vr == 0.0 && vi == 0.0) continue;
    const int g0 = wrapi(it.o[0] + cx, n0), g1 = wrapi(it.o[1] + cy, n1), g2 = wrapi(it.o[2] + cz, n2);
    const long gi = ((long)g0 * n1 + g1) * n2 + g2;
    if (A.c2r) {  // real grid: the real part (the imaginary one is zero for real node values)
      atomicAdd(reinterpret_cast<R*>(A.grid) + gi, (R)vr);
      continue;
    }
    Cplx<R>* cell = A.grid + gi;
    atomicAdd(&cell->x, (R)vr);
    atomicAdd(&cell->y, (R)vi);
  }
}

// ---------------------------------------------------------------------------
// Matrix-core gather (FP32 plans, P = TS + 2m <= 16: tiles of 8 cells for
// m <= 4, of 4 cells for m = 5, 6; f, grad, f + grad).  For a plane a of the tile's padded subgrid and a group of 16
// nodes,
//   T_a[b][j] = sum_c G[a][b][c] Z_j[c]      (Z_j = psi_z of node j, zero-padded)
// is a 16 x 16 x 16 product: four v_mfma_f32_16x16x4f32 (K = 4 values of c
// each) per real and imaginary part, four more each with psi'_z for the
// gradient.  In the accumulator layout a lane holds T_a[b][j] for one node j and
// four rows b, so f_j = sum_a X_j[a] sum_b Y_j[b] T_a[b][j] (and the gradient
// sums, one factor replaced by its derivative) run on the vector ALUs; the four
// lane groups' partial sums over b are added at the end.
//
// One block per nonempty tile (heavy tiles first).  After the subgrid is staged
// in LDS the four waves work independently, each on every fourth group of 16
// sorted nodes: lane group t < 3 of a wave evaluates (or reads from the
// precomputed tables) the axis-t windows of the group's nodes and writes them
// as zero-padded rows into the wave's LDS rows, while the loads for the wave's
// next group are in flight during the matrix-core phase.  A plane that no node
// of the group reaches is skipped by a uniform branch (sorted nodes of a tile
// share their plane offset most of the time).  Nodes found outside their tile
// (bins older than x) are gathered straight from the grid.
// ---------------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// Subgrid rows of 16 floats in LDS, element v of row r at v ^ 4 sw(r),
// sw(r) = (r >> 2) & 3: the float4 chunks k of rows r = 0..15 (one per lane
// group) fall in distinct banks.  Lane group k supplies the MFMA K-slices
// c = 4e + k of the e-th product of a plane from its subgrid chunk (cells in
// the order kp below) and from the z-window row, and reads the rows b = 4k + e
// of its accumulator with the y-window chunk k.
__device__ __forceinline__ int sw4(int r) { return (r >> 1) & 3; }
__device__ __forceinline__ int rpos(int r, int v) { return v ^ (sw4(r) << 2); }

// Position of cell c (axis 2) in a subgrid row: c = 4 e + k at 4 k + e.  The
// float4 chunk k of a row then holds c = k, 4 + k, 8 + k, 12 + k, so the e-th
// product of a plane (component e of every lane's chunk) covers the K-block
// c = 4 e .. 4 e + 3, and a block that no stencil of the group reaches can be
// left out.
__device__ __forceinline__ int kp(int c) { return ((c & 3) << 2) | (c >> 2); }

template <int N>
struct IntC {
  static constexpr int value = N;
};

__device__ __forceinline__ float comp4(const float4& v, int e) {
  return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}

// OR of v over each row of 16 lanes (DPP row shifts); lane 15 of a row holds it
__device__ __forceinline__ unsigned row_or16(unsigned v) {
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  return v;
}

// LDS writes of one lane become visible to the other lanes of its wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NW waves per block: 4, or 2 for tiles of at most 31 nodes (two groups), where
// two more waves would have no group to take
template <int MT, int MODE, int NW = 4>
struct GatherMfmaShape {
  static constexpr int CM = 2 * MT + 1, TS = std_tile(MT, false), P = TS + 2 * MT;
  static constexpr bool WF = (MODE == B_F || MODE == B_FGRAD);
  static constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  // window rows of a wave's 16 nodes: x and z rows of 17 floats (read as
  // scalars), y rows of 20 floats (read as float4 chunks), at X 0, Y YOFF,
  // Z ZOFF; the derivative rows follow at ROWSZ / 2.  The odd and the 20-float
  // pitches keep the rows' stores and reads in distinct banks without a swizzle.
  // With gradients in four-wave blocks (SYZ) the z rows take the y rows' place
  // once those are read (both are read once per group), so such a block fits
  // three times per CU (two-wave blocks fit three times either way).
  static constexpr bool SYZ = WG && NW == 4;
  static constexpr int YOFF = 16 * 17, ZOFF = SYZ ? YOFF : YOFF + 16 * 20;
  static constexpr int ROWSZ = (WG ? 2 : 1) * (SYZ ? YOFF + 16 * 20 : ZOFF + 16 * 17);
  // waves per SIMD the kernel is compiled for (the gradient kernels of m = 5, 6
  // need more than the 168 VGPRs of three)
  static constexpr int WAVES = WG && (MT >= 5 || !SYZ) ? 2 : 3;
  static constexpr int NT = 64 * NW;        // threads per block
  static constexpr size_t LDS = sizeof(float) * (2 * P * 256 + NW * ROWSZ);
};

// one lane's share of a node's window data: axis lk of node li of a group
template <int CM, bool WG>
struct AxisLoad {
  int j;     // node id, -1: no node
  float x;   // coordinate now ...
  float px;  // ... and when the tables were made
  float psi[CM];
  float dpsi[WG ? CM : 1];
};

// blocks take the tiles torder[first + blockIdx.x]
// RG: the real grid of a c2r plan (imaginary parts zero: staged as zeros, and
// the products that would multiply them are left out)
template <int MT, int MODE, int NW, bool RG>
__global__ __launch_bounds__(64 * NW, (GatherMfmaShape<MT, MODE, NW>::WAVES)) void gather_mfma_k(BArgs<float> A,
                                                                                                   int first) {
  using S = GatherMfmaShape<MT, MODE, NW>;
  constexpr int CM = S::CM, TS = S::TS, P = S::P, NT = S::NT;
  constexpr bool WG = S::WG;
  static_assert(P <= 16, "one 16 x 16 block per plane");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* s_gr = reinterpret_cast<float*>(smem_raw);  // [P][16 rows b][16], element c at rpos(b, kp(c))
  float* s_gi = s_gr + P * 256;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  float* rows = s_gi + P * 256 + wave * S::ROWSZ;  // this wave's window rows

  const int tile = A.torder[first + blockIdx.x];
  const int t2 = tile % A.ntile[2];
  const int t1 = (tile / A.ntile[2]) % A.ntile[1];
  const int t0 = tile / (A.ntile[2] * A.ntile[1]);
  const int o0 = t0 * TS, o1 = t1 * TS, o2 = t2 * TS;
  const unsigned n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  const int beg = A.tile_start[tile], end = A.tile_start[tile + 1];

  // stage the subgrid, split into real and imaginary parts; pairs of cells along c
  if constexpr (P < 16) {
    for (int q = tid; q < 2 * P * 256 / 4; q += NT)
      reinterpret_cast<float4*>(s_gr)[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
  }
  {
    constexpr int H = P / 2, NPAIR = P * P * H, K = (NPAIR + NT - 1) / NT;
    const float4* src = reinterpret_cast<const float4*>(A.grid);
    const float2* src_r = reinterpret_cast<const float2*>(A.grid);  // real grid of a c2r plan
    float4 v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = tid + NT * k;
      if (q < NPAIR) {
        const int hz = q % H, cy = (q / H) % P, cx = q / (H * P);
        const unsigned g0 = (unsigned)(o0 + cx) & (n0 - 1), g1 = (unsigned)(o1 + cy) & (n1 - 1);
        const unsigned g2 = (unsigned)(o2 + 2 * hz) & (n2 - 1);
        if constexpr (RG) {
          const float2 r = src_r[((g0 * n1 + g1) * n2 + g2) >> 1];
          v[k] = make_float4(r.x, 0.f, r.y, 0.f);
        } else {
          v[k] = src[((g0 * n1 + g1) * n2 + g2) >> 1];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = tid + NT * k;
      if (q < NPAIR) {
        const int hz = q % H, cy = (q / H) % P, cx = q / (H * P);
        const int r = (cx * 16 + cy) * 16;
        const int p0 = rpos(cy, kp(2 * hz)), p1 = rpos(cy, kp(2 * hz + 1));
        s_gr[r + p0] = v[k].x;
        s_gi[r + p0] = v[k].y;
        s_gr[r + p1] = v[k].z;
        s_gi[r + p1] = v[k].w;
      }
    }
  }
  __syncthreads();

  const int ngroups = (end - beg + 15) >> 4;
  int g = wave;
  if (g >= ngroups) return;  // no block barrier follows

  // this lane's axis (lane group 3 holds none)
  const int ax = lk < 3 ? lk : 0;
  const int nt = ax == 0 ? (int)n0 : (ax == 1 ? (int)n1 : (int)n2);
  const int ot = ax == 0 ? o0 : (ax == 1 ? o1 : o2);
  const float bt = (float)(ax == 0 ? A.g.b[0] : (ax == 1 ? A.g.b[1] : A.g.b[2]));
  const float wst = (float)(ax == 0 ? A.g.ws[0] : (ax == 1 ? A.g.ws[1] : A.g.ws[2]));
  const long M = A.M;
  const bool tab = A.pre_psi != nullptr && (!WG || A.pre_dpsi != nullptr);

  // The loads below are unconditional, from clamped positions (their values
  // are discarded where no node exists), so that no register copy of a loaded
  // value is needed before the matrix-core phase; such a copy would make the
  // wave wait for every load in flight.
  auto perm_raw = [&](int gg) { return A.perm[min(beg + 16 * gg + li, end - 1)]; };
  auto perm_ok = [&](int gg) { return beg + 16 * gg + li < end; };
  auto load_group = [&](int gg, int j, AxisLoad<CM, WG>& L) {
    const long q = min(beg + 16 * gg + li, end - 1);
    L.j = j;
    L.x = A.x[3 * (long)max(j, 0) + ax];
    if (tab) {
      L.px = A.pre_x[(long)ax * M + q];
#pragma unroll
      for (int s = 0; s < CM; s++) L.psi[s] = A.pre_psi[(long)(ax * CM + s) * M + q];
      if (WG)
#pragma unroll
        for (int s = 0; s < CM; s++) L.dpsi[s] = A.pre_dpsi[(long)(ax * CM + s) * M + q];
    }
  };

  AxisLoad<CM, WG> L;
  {
    const int j0 = perm_raw(g);
    load_group(g, perm_ok(g) ? j0 : -1, L);
  }
  int jn = perm_raw(g + NW < ngroups ? g + NW : g);
  bool jn_ok = g + NW < ngroups && perm_ok(g + NW);

  // this lane's window rows (lane group lk < 3: axis lk of node li)
  float* const r0 = rows + (lk == 0 ? 0 : (lk == 1 ? S::YOFF + li * 20 : S::ZOFF)) + (lk == 1 ? 0 : li * 17);
  float* const r1 = r0 + S::ROWSZ / 2;  // the derivative rows (gradient modes)

  for (; g < ngroups; g += NW) {
    const int jg = L.j;
    {
      // the wave's rows are zeroed with consecutive lanes on consecutive 16-byte
      // chunks; the window values below are later stores of the same wave
      const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < (S::ROWSZ / 4 + 63) / 64; k++)
        if (lane + 64 * k < S::ROWSZ / 4) reinterpret_cast<float4*>(rows)[lane + 64 * k] = z4;
    }
    // windows of this lane's axis for node li of the group, written at the
    // node's offset l in the tile (a node outside its tile on this axis leaves
    // the row zero, so it adds nothing; it is listed for the stale kernel below)
    int l = 0;
    bool bad = false;
    // window values of this lane's axis (the z lanes of SYZ write theirs after the y rows are read)
    float zw[S::SYZ ? CM : 1], zdw[S::SYZ ? CM : 1];
    const bool zlate = S::SYZ && lk == 2;
    if (lk < 3 && jg >= 0) {
      const NodeAxis<float> na = node_axis<float>(L.x, nt, MT, A.g.shift);
      l = local_off(na.start, ot, TS, nt);
      bad = l < 0;
      if (tab && L.x == L.px) {
        if constexpr (S::SYZ) {
#pragma unroll
          for (int s = 0; s < CM; s++) {
            zw[s] = L.psi[s];
            zdw[s] = L.dpsi[s];
          }
        }
        if (!bad && !zlate) {
#pragma unroll
          for (int s = 0; s < CM; s++) {
            r0[l + s] = L.psi[s];
            if (WG) r1[l + s] = L.dpsi[s];
          }
        }
      } else {
        float w[CM], dw[CM];
        // inlined: an out-of-line call would keep w and dw in scratch memory
        [[clang::always_inline]] window_axis_s<float, CM, WG>(A.g.win, na.frac, (float)nt, bt, w, dw);
        if constexpr (S::SYZ) {
#pragma unroll
          for (int s = 0; s < CM; s++) {
            zw[s] = w[s] * wst;
            zdw[s] = dw[s] * wst;
          }
        }
        if (!bad && !zlate) {
#pragma unroll
          for (int s = 0; s < CM; s++) {
            r0[l + s] = w[s] * wst;
            if (WG) r1[l + s] = dw[s] * wst;
          }
        }
      }
    }
    const unsigned long long badm = __ballot(bad);
    const unsigned stale = (unsigned)(badm | (badm >> 16) | (badm >> 32)) & 0xFFFFu;
    const bool ok = jg >= 0 && !((stale >> li) & 1u);
    // the offsets of the group's nodes per axis, as bit masks: an OR over the
    // 16 lanes of each lane group (lane 15 of the x group, lane 47 of the z group)
    const unsigned om = row_or16((lk < 3 && ok) ? (1u << l) : 0u);
    const unsigned xm = (unsigned)__builtin_amdgcn_readlane((int)om, 15);
    const unsigned zm = (unsigned)__builtin_amdgcn_readlane((int)om, 47);
    // planes the group's stencils reach, [amin, amax]
    const int amin = xm ? __builtin_ctz(xm) : P, amax = xm ? 31 - __builtin_clz(xm) + CM - 1 : -1;
    wave_sync();

    // loads for the wave's next group, in flight during the matrix-core phase
    AxisLoad<CM, WG> Ln;
    const int gn = g + NW < ngroups ? g + NW : g;
    load_group(gn, jn_ok ? jn : -1, Ln);
    const int jnn = perm_raw(g + 2 * NW < ngroups ? g + 2 * NW : g);
    const bool jnn_ok = g + 2 * NW < ngroups && perm_ok(g + 2 * NW);

    const float* xr = rows + li * 17;
    const float* dxr = xr + S::ROWSZ / 2;
    const int ch = 4 * (lk ^ sw4(li));
    const float4 yb = *reinterpret_cast<const float4*>(rows + S::YOFF + li * 20 + 4 * lk);  // Y_j[4 lk + e]
    float4 dyb = yb;
    if (WG) dyb = *reinterpret_cast<const float4*>(rows + S::ROWSZ / 2 + S::YOFF + li * 20 + 4 * lk);
    if constexpr (S::SYZ) {
      // the y rows are read: clear their area (values and derivatives) and write the z rows there
      wave_sync();
      const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
      constexpr int NQ = 16 * 20 / 4;  // 16-byte chunks of one area
#pragma unroll
      for (int k = 0; k < (2 * NQ + 63) / 64; k++) {
        const int qq = lane + 64 * k;
        if (qq < 2 * NQ)
          reinterpret_cast<float4*>(rows + S::YOFF + (qq >= NQ ? S::ROWSZ / 2 - NQ * 4 : 0))[qq] = z4;
      }
      if (zlate && jg >= 0 && !bad) {
        float* const rz = rows + S::ZOFF + li * 17;
#pragma unroll
        for (int s = 0; s < CM; s++) {
          rz[l + s] = zw[s];
          rz[S::ROWSZ / 2 + l + s] = zdw[s];
        }
      }
      wave_sync();
    }
    // Z_j[4 e + lk] in component e: the K-slice of this lane group in the e-th product
    auto zread = [&](const float* row) { return make_float4(row[lk], row[4 + lk], row[8 + lk], row[12 + lk]); };
    const float4 zc = zread(rows + S::ZOFF + li * 17);
    float4 dzc = zc;
    if (WG) dzc = zread(rows + S::ROWSZ / 2 + S::ZOFF + li * 17);
    float fr = 0.f, fi = 0.f, g0r = 0.f, g0i = 0.f, g1r = 0.f, g1i = 0.f, g2r = 0.f, g2i = 0.f;
    // plane a + 1's operands are read while plane a is on the matrix cores;
    // products e = K0 .. K0 + NK - 1 (the K-blocks the group's stencils reach)
    auto plane_loop = [&](auto k0c, auto nkc) __attribute__((always_inline)) {
      constexpr int K0 = decltype(k0c)::value, NK = decltype(nkc)::value;
      auto ld_plane = [&](int a, float4& gr, float4& gi, float& xa, float& dxa) {
        gr = *reinterpret_cast<const float4*>(s_gr + (a * 16 + li) * 16 + ch);
        gi = *reinterpret_cast<const float4*>(s_gi + (a * 16 + li) * 16 + ch);
        xa = xr[a];
        dxa = WG ? dxr[a] : 0.f;
      };
      float4 gr, gi;
      float xa, dxa;
      ld_plane(amax >= amin ? amin : 0, gr, gi, xa, dxa);
      // sums over all planes with X_j[a] folded into the products' second operand:
      // T = sum_a sum_c G[a][b][c] X_j[a] Z_j[c] (f mode), or with Z' (the z
      // gradient in the gradient modes); their y sums follow after the loop
      mfma_f4 vr = {0.f, 0.f, 0.f, 0.f}, vi = vr;
      for (int a = amin; a <= amax; a++) {
        float4 grn, gin;
        float xan, dxan;
        ld_plane(a < amax ? a + 1 : a, grn, gin, xan, dxan);
        if (!WG) {
#pragma unroll
          for (int e = K0; e < K0 + NK; e++) {
            const float bz = xa * comp4(zc, e);
            vr = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gr, e), bz, vr, 0, 0, 0);
            if constexpr (!RG) vi = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gi, e), bz, vi, 0, 0, 0);
          }
        } else {
          mfma_f4 tr = {0.f, 0.f, 0.f, 0.f}, ti = tr;
#pragma unroll
          for (int e = K0; e < K0 + NK; e++) {
            tr = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gr, e), comp4(zc, e), tr, 0, 0, 0);
            if constexpr (!RG) ti = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gi, e), comp4(zc, e), ti, 0, 0, 0);
          }
#pragma unroll
          for (int e = K0; e < K0 + NK; e++) {
            const float bz = xa * comp4(dzc, e);
            vr = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gr, e), bz, vr, 0, 0, 0);
            if constexpr (!RG) vi = __builtin_amdgcn_mfma_f32_16x16x4f32(comp4(gi, e), bz, vi, 0, 0, 0);
          }
          const float sr = yb.x * tr[0] + yb.y * tr[1] + yb.z * tr[2] + yb.w * tr[3];
          const float si = yb.x * ti[0] + yb.y * ti[1] + yb.z * ti[2] + yb.w * ti[3];
          if (S::WF) {
            fr += xa * sr;
            fi += xa * si;
          }
          const float syr = dyb.x * tr[0] + dyb.y * tr[1] + dyb.z * tr[2] + dyb.w * tr[3];
          const float syi = dyb.x * ti[0] + dyb.y * ti[1] + dyb.z * ti[2] + dyb.w * ti[3];
          g0r += dxa * sr;
          g0i += dxa * si;
          g1r += xa * syr;
          g1i += xa * syi;
        }
        gr = grn;
        gi = gin;
        xa = xan;
        dxa = dxan;
      }
      const float vsr = yb.x * vr[0] + yb.y * vr[1] + yb.z * vr[2] + yb.w * vr[3];
      const float vsi = yb.x * vi[0] + yb.y * vi[1] + yb.z * vi[2] + yb.w * vi[3];
      if (!WG) {
        fr = vsr;
        fi = vsi;
      } else {
        g2r = vsr;
        g2i = vsi;
      }
    };
    if constexpr (MT == 4 && P == 16) {
      // a 9-point stencil starting at l reaches the K-blocks l / 4 .. l / 4 + 2; the
      // group needs all four only when its z offsets lie on both sides of 4
      const bool lo = (zm & 0x0Fu) != 0, hi = (zm & 0xF0u) != 0;
      if (lo && hi) plane_loop(IntC<0>{}, IntC<4>{});
      else if (lo) plane_loop(IntC<0>{}, IntC<3>{});
      else plane_loop(IntC<1>{}, IntC<3>{});
    } else {
      plane_loop(IntC<0>{}, IntC<4>{});
    }
    // sum over the four lane groups (rows b = 4 lk + e)
    auto red = [](float v) {
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      return v;
    };
    fr = red(fr);
    fi = red(fi);
    if (WG) {
      g0r = red(g0r);
      g0i = red(g0i);
      g1r = red(g1r);
      g1i = red(g1i);
      g2r = red(g2r);
      g2i = red(g2i);
    }
    if (lk == 0 && jg >= 0) {
      if (ok) {
        Cplx<float> v[4] = {cmk<float>(fr, fi), cmk<float>(g0r, g0i), cmk<float>(g1r, g1i), cmk<float>(g2r, g2i)};
        gather_store<float, MODE>(A, jg, v);
      } else {
        A.stale_list[atomicAdd(A.stale_count, 1u)] = beg + 16 * g + li;  // gathered by gather_stale_k
      }
    }
    L = Ln;
    if (g + NW >= ngroups) L.j = -1;
    jn = jnn;
    jn_ok = jnn_ok;
    wave_sync();
  }
}

__global__ void zero_u32_k(unsigned* c) {
  if (threadIdx.x == 0) *c = 0u;
}

// nodes the matrix-core gather listed as outside their tile: gathered straight
// from the grid
template <int MT, int MODE>
__global__ __launch_bounds__(256) void gather_stale_k(BArgs<float> A) {
  constexpr int CM = 2 * MT + 1;
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  const unsigned count = *A.stale_count;
  for (unsigned k = blockIdx.x * 256 + threadIdx.x; k < count; k += gridDim.x * 256) {
    const long j = A.perm[A.stale_list[k]];
    NodeWin<float, CM> w;
    node_windows<float, MT, CM, WG>(A, -1, j, MT, w);
    gather_node_global<float, CM, MODE, true>(A, j, w, CM);
  }
}

// Tiles of one node (the last n_single of the tile order): one wave per node,
// gathered straight from the grid (the matrix-core kernels would stage a whole
// subgrid for it).  Every lane forms the node's windows (the tables when they
// hold its coordinates), lane r takes the stencil rows (a, b) = (r / CM, r % CM)
// and r + 64, and the wave adds the lanes' partial sums.
template <int MT, int MODE>
__global__ __launch_bounds__(256) void gather_single_k(BArgs<float> A, int first) {
  constexpr int CM = 2 * MT + 1, NR = CM * CM;
  constexpr bool WG = (MODE == B_GRAD || MODE == B_FGRAD);
  const int lane = threadIdx.x & 63;
  const int p = first + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= A.n_torder) return;  // whole wave; no barrier follows
  const int pos = A.tile_start[A.torder[p]];  // the tile's node, as a sorted position
  const long j = A.perm[pos];
  NodeWin<float, CM> w;
  node_windows<float, MT, CM, WG>(A, pos, j, MT, w);
  const int n0 = A.g.n[0], n1 = A.g.n[1], n2 = A.g.n[2];
  int i2s[CM];
#pragma unroll
  for (int s = 0; s < CM; s++) {
    const int v = w.start[2] + s;
    i2s[s] = v >= n2 ? v - n2 : v;
  }
  float fr = 0.f, fi = 0.f, g0r = 0.f, g0i = 0.f, g1r = 0.f, g1i = 0.f, g2r = 0.f, g2i = 0.f;
#pragma unroll
  for (int it = 0; it < (NR + 63) / 64; it++) {
    const int r = lane + 64 * it;
    const bool live = r < NR;
    const int a = live ? r / CM : 0, b = live ? r - (r / CM) * CM : 0;
    int i0 = w.start[0] + a;
    if (i0 >= n0) i0 -= n0;
    int i1 = w.start[1] + b;
    if (i1 >= n1) i1 -= n1;
    const long off = ((long)i0 * n1 + i1) * n2;
    float hr = 0.f, hi = 0.f, hzr = 0.f, hzi = 0.f;
    if (A.c2r) {  // real grid of a c2r plan
      const float* row = reinterpret_cast<const float*>(A.grid) + off;
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const float gv = row[i2s[s]];
        hr += w.pz[s] * gv;
        if (WG) hzr += w.dz[s] * gv;
      }
    } else {
      const Cplx<float>* row = A.grid + off;
#pragma unroll
      for (int s = 0; s < CM; s++) {
        const Cplx<float> gv = row[i2s[s]];
        hr += w.pz[s] * gv.x;
        hi += w.pz[s] * gv.y;
        if (WG) {
          hzr += w.dz[s] * gv.x;
          hzi += w.dz[s] * gv.y;
        }
      }
    }
    // the windows of plane a and row b (lane-dependent indices, picked by selects)
    float xa = w.px[0], yb = w.py[0], dxa = WG ? w.dx[0] : 0.f, dyb = WG ? w.dy[0] : 0.f;
#pragma unroll
    for (int k = 1; k < CM; k++) {
      xa = a == k ? w.px[k] : xa;
      yb = b == k ? w.py[k] : yb;
      if (WG) {
        dxa = a == k ? w.dx[k] : dxa;
        dyb = b == k ? w.dy[k] : dyb;
      }
    }
    const float xy = live ? xa * yb : 0.f;
    fr += xy * hr;
    fi += xy * hi;
    if (WG) {
      const float dxy = live ? dxa * yb : 0.f, xdy = live ? xa * dyb : 0.f;
      g0r += dxy * hr;
      g0i += dxy * hi;
      g1r += xdy * hr;
      g1i += xdy * hi;
      g2r += xy * hzr;
      g2i += xy * hzi;
    }
  }
  auto red = [](float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
  };
  Cplx<float> v[4] = {cmk<float>(red(fr), red(fi)), cmk<float>(0.f, 0.f), cmk<float>(0.f, 0.f),
                      cmk<float>(0.f, 0.f)};
  if (WG) {
    v[1] = cmk<float>(red(g0r), red(g0i));
    v[2] = cmk<float>(red(g1r), red(g1i));
    v[3] = cmk<float>(red(g2r), red(g2i));
  }
  if (lane == 0) gather_store<float, MODE>(A, j, v);
}

inline unsigned grid1d(long n, int t) { return (unsigned)((n + t - 1) / t); }

// the matrix-core gather applies to FP32 f / grad / f + grad with m <= 6 on the
// standard tile layout (padded subgrids of at most 16 cells per axis), once the
// nonempty tiles are listed
template <class R, int MT, int MODE>
bool mfma_gather(const BArgs<R>& A) {
  if constexpr (sizeof(R) != 4 || MT > 6 || MODE > B_FGRAD) {
    return false;
  } else {
    constexpr int TS = GatherMfmaShape<MT, MODE, 4>::TS;
    return !A.no_mfma && A.torder && A.n_torder > 0 && A.stale_count && A.stale_list && A.ts[0] == TS &&
           A.ts[1] == TS && A.ts[2] == TS;
  }
}

template <class R, int MT, int MODE>
void launch_gather_mfma(const BArgs<R>& A, hipStream_t s) {
  if constexpr (sizeof(R) == 4 && MT <= 6 && MODE <= B_FGRAD) {
    using S4 = GatherMfmaShape<MT, MODE, 4>;
    using S2 = GatherMfmaShape<MT, MODE, 2>;
    const int nh = min(max(A.n_heavy, 0), A.n_torder);
    const int nl = max(nh, A.n_torder - max(A.n_single, 0));  // [nl, n_torder): tiles of one node
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, A.stale_count);
    auto go = [&](auto rgc) {
      constexpr bool RG = decltype(rgc)::value;
      if (nh > 0) {
        (void)hipFuncSetAttribute((const void*)gather_mfma_k<MT, MODE, 4, RG>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)S4::LDS);
        hipLaunchKernelGGL((gather_mfma_k<MT, MODE, 4, RG>), dim3((unsigned)nh), dim3(S4::NT), S4::LDS, s, A, 0);
      }
      if (nl > nh) {  // tiles of 2..31 nodes: two groups, two waves
        (void)hipFuncSetAttribute((const void*)gather_mfma_k<MT, MODE, 2, RG>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)S2::LDS);
        hipLaunchKernelGGL((gather_mfma_k<MT, MODE, 2, RG>), dim3((unsigned)(nl - nh)), dim3(S2::NT),
                           S2::LDS, s, A, nh);
      }
    };
    if (A.c2r)
      go(std::true_type{});
    else
      go(std::false_type{});
    if (A.n_torder > nl)
      hipLaunchKernelGGL((gather_single_k<MT, MODE>), dim3(grid1d(A.n_torder - nl, 4)), dim3(256), 0, s, A, nl);
    if (A.mid) (void)hipEventRecord(A.mid, s);
    if (A.mid2) (void)hipEventRecord(A.mid2, s);
    hipLaunchKernelGGL((gather_stale_k<MT, MODE>), dim3(256), dim3(256), 0, s, A);
  }
}

}  // namespace

template <class R, int KIND, int MODE>
int b_launch(const BArgs<R>& A, hipStream_t s) {
  constexpr int SMODE = MODE > 3 ? 3 : MODE;
  const dim3 grid(grid1d(A.M, kB)), block(kB);
  const bool tiled = A.items != nullptr && A.n_items > 0 && A.g.m >= 2 && A.g.m <= 8;
  size_t lds = 0;
  if (tiled) {
    // padded subgrid: P0 planes of pitch P1*(P2+1)+1 cells (SubgridPitch)
    const long P0 = A.ts[0] + 2 * A.g.m, P1 = A.ts[1] + 2 * A.g.m, P2 = A.ts[2] + 2 * A.g.m;
    const long cells = P0 * (P1 * (P2 + 1) + 1);
    lds = (size_t)cells * (KIND == 0 ? sizeof(Cplx<R>) : 16);
  }
  const dim3 tgrid((unsigned)(tiled ? A.n_items : 1)), tblock(kTileThreads);
  BArgs<R> D = A;
  if (!tiled) D.tile_dense = nullptr;  // no tiled pass: the direct kernel takes every node
  switch (A.g.m) {
#define PNFFT_B_CASE(MV)                                                                                  \
  case MV:                                                                                                \
    if constexpr (KIND == 0) {                                                                            \
      if (mfma_gather<R, MV, MODE>(A)) {                                                                  \
        launch_gather_mfma<R, MV, MODE>(A, s);                                                            \
        break;                                                                                            \
      }                                                                                                   \
      if (tiled) {                                                                                        \
        (void)hipFuncSetAttribute((const void*)gather_tiled_k<R, MV, MODE>,                               \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                  \
        hipLaunchKernelGGL((gather_tiled_k<R, MV, MODE>), tgrid, tblock, lds, s, A);                      \
      }                                                                                                   \
      if (A.mid) (void)hipEventRecord(A.mid, s);                                                          \
      if (A.mid2) (void)hipEventRecord(A.mid2, s);                                                        \
      if (!(tiled && A.all_dense)) hipLaunchKernelGGL((gather_direct_k<R, MV, MODE>), grid, block, 0, s, D);  \
    } else {                                                                                              \
      if (tiled) {                                                                                        \
        (void)hipFuncSetAttribute((const void*)scatter_tiled_k<R, MV, SMODE>,                             \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                  \
        hipLaunchKernelGGL((scatter_tiled_k<R, MV, SMODE>), tgrid, tblock, lds, s, A);                    \
      }                                                                                                   \
      if (A.mid) (void)hipEventRecord(A.mid, s);                                                          \
      if (A.mid2) (void)hipEventRecord(A.mid2, s);                                                        \
      if (!(tiled && A.all_dense)) hipLaunchKernelGGL((scatter_direct_k<R, MV, SMODE>), grid, block, 0, s, D); \
    }                                                                                                     \
    break;
    PNFFT_B_CASE(2)
    PNFFT_B_CASE(3)
    PNFFT_B_CASE(4)
    PNFFT_B_CASE(5)
    PNFFT_B_CASE(6)
    PNFFT_B_CASE(7)
    PNFFT_B_CASE(8)
#undef PNFFT_B_CASE
    default:  // plans are limited to m = 2..8
      return kUnsupported;
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template int b_launch<PNFFT_B_REAL, PNFFT_B_KIND, PNFFT_B_MODE>(const BArgs<PNFFT_B_REAL>&, hipStream_t);

}  // namespace pnfft_amd
