// nfft_b.hpp -- host interface of matrix B (gather / scatter) and node binning.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "window_ext.hpp"

namespace pnfft_amd {

// B modes.  Gather: f, grad, f+grad, or the x-VJP epilogue.  Scatter: f, grad, f+grad.
enum BMode : int { B_F = 1, B_GRAD = 2, B_FGRAD = 3, B_VJP = 4 };

template <class R>
struct BArgs {
  GridGeom g;
  long M;
  const R* x;              // R[3M]
  const int* perm;         // processing order (sorted node ids) or null
  Cplx<R>* grid;           // n0*n1*n2, natural order
  Cplx<R>* f;              // C[M]
  Cplx<R>* grad;           // C[3M]
  const Cplx<R>* w;        // VJP cotangent C[M]
  R* gx;                   // VJP output R[3M]
  int accumulate;          // gather: add onto outputs
  int real_only;           // gather: PNFFT_REAL_F, write the real part only
  int c2r;                 // real plan: f is R[M], grad R[3M], w R[M]; the real grid part is used
  // tiled path
  const int* tile_start;   // [ntiles+1] offsets into perm
  const unsigned* keys;    // sorted bin keys [M]
  const unsigned char* tile_dense;  // [ntiles] 1 if the tile's nodes go through the LDS path
  const int* items;        // work items: pairs (tile id, chunk index)
  int n_items;
  int ts[3];               // tile size per axis
  int ntile[3];            // tiles per axis
  int all_dense;           // every nonempty tile is a work item: the direct kernel has no nodes to take
  // nonempty tiles, most nodes first (column scatter and matrix-core gather)
  const int* torder;       // [n_torder]
  int n_torder;
  int n_heavy;             // tiles torder[0, n_heavy) hold 32 nodes or more, the rest at most 31
  int n_single;            // the last n_single tiles of torder hold one node each
  // column scatter (scatter_cols.hip): when slots is set, scatter() runs it instead of the LDS-atomic path
  Cplx<R>* slots;          // [n_torder][P^3] padded subgrid of every nonempty tile
  const int* tile_slot;    // [ntiles] slot of each tile, -1 if empty
  // offsets of each tile's nodes at binning time, per axis (bits 8 t + l for
  // offset l on axis t), by tile id and by position in torder (or null)
  const unsigned* box_tile;
  const unsigned* box_ord;
  int box_slots;           // scatter: slots hold only the cells of their box (the fold reads box_tile)
  int no_mfma;             // 1: vector kernels also where the matrix-core gather / scatter apply
  // window tables of pnfft_precompute_psi in the sorted order, [3 CM][M] each (or null)
  const R* pre_psi;
  const R* pre_dpsi;
  const R* pre_x;          // [3][M] coordinates the tables were computed from
  unsigned* stale_count;   // nodes found outside the tile they are binned to (bins older than x) ...
  int* stale_list;         // ... and their sorted positions [M]
  hipEvent_t mid;          // host only: recorded between the first and the second kernel of the stage
  hipEvent_t mid2;         // host only: recorded between the second and the third kernel of the stage
};

// Tile size (cells per axis) of the standard binning layout: 8 for m <= 4 and
// FP64 m = 5, 6; 4 for FP32 m = 5, 6 (padded subgrids of 4 + 2m <= 16 cells,
// the matrix-core kernels' shape) and for m >= 7 (padded subgrids of FP64
// accumulators within the LDS).
constexpr int std_tile(int m, bool is_double) { return (m <= 4 || (m <= 6 && is_double)) ? 8 : 4; }

constexpr int kTileThreads = 512;  // threads per block of the tiled kernels
constexpr int kTileNodes = 512;    // nodes per work item

// Column scatter: applies for m = 2..8 (FP64: m <= 6) when all three tile
// sizes equal the layout's TS (std_tile); P = TS + 2m cells per
// axis of a tile's padded subgrid.  Returns P, or 0 if the path does not apply.
int cols_pad(int m, bool is_double, const int* ts);

// Nonempty tiles: those of 32 nodes or more by decreasing node count
// (power-of-two classes, most nodes first), then those of 2..31 nodes and then
// those of one node, both in tile order; tile_slot[t] = position of tile t (-1
// if empty).  work holds 65 unsigned, partial the scan partials of ntiles values.
// Returns the number of nonempty tiles (host value; synchronizes s), or -1;
// *n_heavy = the number of them with 32 nodes or more; the tiles of one node
// come last, *n_single of them.  light holds 2 ntiles unsigned.
int build_tile_order(const int* tile_start, long ntiles, unsigned* work, unsigned* light, unsigned* partial,
                     int* order, int* tile_slot, int* n_heavy, int* n_single, hipStream_t s);

// Launch of the column scatter for MODE (B_F, B_GRAD, B_FGRAD): column kernel
// into the slots, fold of the slots into the grid (which it overwrites), then
// the nodes the column kernel listed as outside their tile (bins older than
// their coordinates) through the direct path (global atomics).
template <class R, int MODE>
int cols_launch(const BArgs<R>& A, hipStream_t s);

// Work items of the tiled kernels: every tile with at least dense_min nodes is
// cut into chunks of kTileNodes sorted nodes.  Writes tile_dense[ntiles],
// items[2 * n_items] and returns n_items (host value; synchronizes s), or -1.
int build_items(const int* tile_start, long ntiles, int dense_min, unsigned* chunks, unsigned* partial,
                unsigned char* tile_dense, int* items, long items_cap, hipStream_t s);

// Tile layout used by the binning: a node belongs to the tile that holds its
// stencil start (floor(n x) - m) mod n on every axis.
struct TileLayout {
  int ts[3];
  int ntile[3];
  long ntiles() const { return (long)ntile[0] * ntile[1] * ntile[2]; }
  long tile_vol() const { return (long)ts[0] * ts[1] * ts[2]; }
};

// Per-tile offset masks of the nodes whose sorted bin keys are keys[0, M)
// (tile layout T, tile sizes of at most 8): box_tile[ntiles] by tile id and
// box_ord[n_order] by position in order.  The slots of the matrix-core column
// scatter (with the brick fold) hold only the box these masks span (from the
// smallest offset to the largest plus 2m per axis); a node whose stencil
// leaves that box (moved since binning) takes the stale path.  Returns kOk,
// kUnsupported for larger tiles, or kHipError.
int build_tile_boxes(const unsigned* keys, long M, const TileLayout& T, const int* order, int n_order,
                     unsigned* box_tile, unsigned* box_ord, hipStream_t s);

// Sorting scratch (device): keys/values double-buffered, histograms, scan partials.
struct SortWork {
  unsigned* keys[2] = {nullptr, nullptr};
  int* vals[2] = {nullptr, nullptr};
  unsigned* hist = nullptr;      // 256 * nblocks
  unsigned* partial = nullptr;   // scan partials
  long cap = 0;                  // capacity in nodes
  long hist_cap = 0;
  long partial_cap = 0;
};

// Compute bin keys of every node and sort the node ids by key (stable LSD
// radix sort, 8-bit digits).  Result: perm (sorted node ids, == vals[res]),
// sorted keys, and tile_start[ntiles+1].  Returns kOk or an error code.
template <class R>
int bin_nodes(const R* x, long M, const GridGeom& g, const TileLayout& T, SortWork& W, int* perm_out,
              unsigned* keys_out, int* tile_start, hipStream_t s);

// Kernel launchers, one object file per (R, KIND, MODE); KIND 0 = gather, 1 = scatter.
template <class R, int KIND, int MODE>
int b_launch(const BArgs<R>& A, hipStream_t s);

template <class R>
int gather(const BArgs<R>& A, int mode, hipStream_t s);

// grid = B^T (f, grad): the whole grid is written (the column path overwrites
// it through the fold; the LDS-atomic path clears it first).
template <class R>
int scatter(const BArgs<R>& A, int mode, hipStream_t s);

template <class R>
int fill_zero(Cplx<R>* p, size_t n, hipStream_t s);

// Window tables for pnfft_precompute_psi: for sorted position p (node perm[p]),
// psi[(t CM + s) M + p] = psi_t(s) (and dpsi likewise if non-null), x_out[t M + p]
// = x_t.  Applies for m = 2..8 (FP64: m <= 6); returns kUnsupported otherwise.
template <class R>
int precompute_tables(const BArgs<R>& A, R* psi, R* dpsi, R* x_out, hipStream_t s);

// Binning that allocates nothing: the sort's scratch is a file-scope device
// array (noalloc_scratch_bytes() bytes per GPU, held until the process ends)
// sized for kNoAllocMaxNodes nodes.  Writes perm_out / keys_out (device arrays
// of M entries) as bin_nodes does; no tile bounds.  Returns kUnsupported for
// M > kNoAllocMaxNodes, kHipError if the array's address cannot be resolved.
// Calls that overlap on one GPU share the array and race.
constexpr long kNoAllocMaxNodes = 1L << 20;
template <class R>
int bin_nodes_noalloc(const R* x, long M, const GridGeom& g, const TileLayout& T, int* perm_out, unsigned* keys_out,
                      hipStream_t s);
size_t noalloc_scratch_bytes();

// Hessian gather (PNFFT_COMPUTE_HESSIAN_F by window derivatives): tab holds
// 9 (2m+1) M reals and st 3 M ints of scratch (the node's psi, psi', psi'' per
// axis, from I's tables for interpolated windows, and stencil starts);
// hess[6j + q], q = xx, xy, xz, yy, yz, zz, is written (added onto when
// A.accumulate; R[6M] for c2r plans).  m = 2..8.
template <class R>
int hessian_gather(const BArgs<R>& A, const IntpolTabs<R>& I, Cplx<R>* hess, R* tab, int* st, hipStream_t s);

// Window tables of the Bessel-I0 window and of interpolated windows (the
// layout of precompute_tables: psi[(t CM + s) M + p], dpsi likewise if
// non-null, x_out[t M + p]), for the nodes in A.perm's order and A.g.
template <class R>
int ext_tables(const BArgs<R>& A, const IntpolTabs<R>& I, R* psi, R* dpsi, R* x_out, hipStream_t s);

// Strided node-value copies of the ik-differentiated transforms (comps = 2 for
// complex values, 1 for the reals of c2r plans):
//   put: out[(j stride + offset) comps + c] = (acc ? out[..] : 0) + in[j comps + c]
//   get: out[j comps + c] = in[(j stride + offset) comps + c]
template <class R>
int strided_put(R* out, const R* in, long M, int comps, int stride, int offset, int acc, hipStream_t s);
template <class R>
int strided_get(R* out, const R* in, long M, int comps, int stride, int offset, hipStream_t s);

// Number of sort passes for keys below key_limit.
int sort_passes(unsigned long long key_limit);

}  // namespace pnfft_amd
