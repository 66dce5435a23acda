// nfft_b.hip -- matrix B of the NFFT (window gather for trafo, window scatter
// for adj) and the binning of nodes by grid cell.
//
// Reference: PNX(trafo_B_ad) / PNX(adjoint_B_ad) and their node loops
// (kernel/ndft-parallel.c:2521-3009) with the stencil kernels of
// kernel/assign.c:478-1066.  The reference stores the FFT output shifted and
// ghost-padded by m cells and exchanges the ghost layers (PX(exchange) /
// PX(reduce), :2558, :2679); here the grid stays in natural DFT order and the
// periodic wrap is applied to the stencil indices, so no ghost copy exists.
//
// Node n's stencil on axis t covers grid indices (floor(n_t x_t) - m + s) mod n_t,
// s = 0..2m (cutoff 2m+1, :910, :2165-2174).  floor(n x) is taken in FP64 for
// both precisions; the window is evaluated in the compute precision from the
// FP64 fractional part.
#include <cstdio>
#include <mutex>

#include "device_common.hpp"
#include "b_common.hpp"
#include "nfft_b.hpp"
#include "window_math.hpp"

namespace pnfft_amd {

namespace {

constexpr int kB = 256;  // threads per block of the B kernels

// ---------------------------------------------------------------------------
// binning: keys, stable LSD radix sort, tile bounds
// ---------------------------------------------------------------------------
constexpr int kSortT = 256;
constexpr int kSortIPB = 4096;  // items per block
constexpr int kScanChunk = 4096;

template <class R>
__global__ __launch_bounds__(kB) void bin_keys_k(const R* __restrict__ x, long M, GridGeom g, TileLayout T,
                                                 unsigned* keys, int* vals) {
  const long j = (long)blockIdx.x * kB + threadIdx.x;
  if (j >= M) return;
  unsigned long long tile = 0;
  unsigned l[3];
  for (int t = 0; t < 3; t++) {
    const NodeAxis<R> a = node_axis<R>(x[3 * j + t], g.n[t], g.m, g.shift);
    tile = tile * T.ntile[t] + (unsigned)(a.start / T.ts[t]);
    l[t] = (unsigned)(a.start % T.ts[t]);
  }
  // cell order inside a tile: runs of q = min(TS2, 4) cells along axis 2
  // outermost, then axis 0, axis 1 and the cell within the run, so that
  // consecutive nodes of a tile share the quarter of axis 2 their stencils start
  // in (the matrix-core gather then skips one of its four K-blocks for most
  // 16-node groups)
  const unsigned q = T.ts[2] < 4 ? (unsigned)T.ts[2] : 4u;
  const unsigned long long in =
      ((unsigned long long)(l[2] / q) * (T.ts[0] * T.ts[1]) + l[0] * T.ts[1] + l[1]) * q + l[2] % q;
  keys[j] = (unsigned)(tile * (unsigned long long)(T.ts[0] * T.ts[1] * T.ts[2]) + in);
  vals[j] = (int)j;
}

__global__ __launch_bounds__(kSortT) void sort_hist_k(const unsigned* __restrict__ keys, long M, int shift,
                                                      unsigned* hist, int nblk) {
  __shared__ unsigned h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const long base = (long)blockIdx.x * kSortIPB;
  for (int i = 0; i < kSortIPB / kSortT; i++) {
    const long idx = base + (long)i * kSortT + tid;
    if (idx < M) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(long)tid * nblk + blockIdx.x] = h[tid];
}

// Block-wide exclusive scan of 256 values held one per thread.
__device__ inline unsigned block_exscan256(unsigned v, unsigned* tmp, unsigned* total) {
  const int tid = threadIdx.x;
  tmp[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    unsigned add = (tid >= off) ? tmp[tid - off] : 0u;
    __syncthreads();
    tmp[tid] += add;
    __syncthreads();
  }
  const unsigned incl = tmp[tid];
  if (total) *total = tmp[255];
  __syncthreads();
  return incl - v;
}

// chunk sums
__global__ __launch_bounds__(kSortT) void scan_reduce_k(const unsigned* __restrict__ d, long n, unsigned* partial) {
  __shared__ unsigned tmp[256];
  const long base = (long)blockIdx.x * kScanChunk;
  unsigned s = 0;
  for (int i = 0; i < kScanChunk / kSortT; i++) {
    const long idx = base + (long)threadIdx.x * (kScanChunk / kSortT) + i;
    if (idx < n) s += d[idx];
  }
  unsigned total;
  __shared__ unsigned tot;
  (void)block_exscan256(s, tmp, &tot);
  total = tot;
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// exclusive scan of the chunk sums, one block, sequential over 256-wide pieces
__global__ __launch_bounds__(kSortT) void scan_partials_k(unsigned* partial, long np) {
  __shared__ unsigned tmp[256];
  __shared__ unsigned tot;
  unsigned carry = 0;
  for (long base = 0; base < np; base += 256) {
    const long idx = base + threadIdx.x;
    const unsigned v = idx < np ? partial[idx] : 0u;
    const unsigned ex = block_exscan256(v, tmp, &tot);
    if (idx < np) partial[idx] = ex + carry;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSortT) void scan_apply_k(unsigned* d, long n, const unsigned* __restrict__ partial) {
  __shared__ unsigned tmp[256];
  __shared__ unsigned tot;
  constexpr int PER = kScanChunk / kSortT;
  const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * PER;
  unsigned loc[PER];
  unsigned s = 0;
  for (int i = 0; i < PER; i++) {
    const long idx = base + i;
    loc[i] = idx < n ? d[idx] : 0u;
    s += loc[i];
  }
  unsigned ex = block_exscan256(s, tmp, &tot) + partial[blockIdx.x];
  for (int i = 0; i < PER; i++) {
    const long idx = base + i;
    if (idx < n) d[idx] = ex;
    ex += loc[i];
  }
}

// Stable scatter of one digit pass: items of a block are ranked in index order
// and placed by digit in LDS first, so that the stores to the output are runs
// of consecutive positions (one per digit present in the block) rather than
// one scattered 4-byte store per item.  offs: the exclusive scan of the
// per-(digit, block) counts in digit-major order, so that a count is the
// difference of neighbouring entries.
__global__ __launch_bounds__(kSortT) void sort_scatter_k(const unsigned* __restrict__ kin, const int* __restrict__ vin,
                                                         unsigned* kout, int* vout, long M, int shift,
                                                         const unsigned* __restrict__ offs, int nblk) {
  __shared__ unsigned run[256];
  __shared__ unsigned wcnt[4][256];
  __shared__ unsigned loc[256];   // first LDS position of digit d
  __shared__ unsigned gdst[256];  // output position of LDS position 0 for digit d
  __shared__ unsigned tmp[256];
  __shared__ unsigned sk[kSortIPB];
  __shared__ int sv[kSortIPB];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const long base = (long)blockIdx.x * kSortIPB;
  {
    const long e = (long)tid * nblk + blockIdx.x;  // entry (digit tid, this block)
    const unsigned o = offs[e];
    const unsigned c = (e + 1 < 256L * nblk ? offs[e + 1] : (unsigned)M) - o;
    const unsigned l = block_exscan256(c, tmp, nullptr);
    loc[tid] = l;
    gdst[tid] = o - l;
    run[tid] = 0;
  }
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortIPB / kSortT; r++) {
    const long idx = base + (long)r * kSortT + tid;
    const bool valid = idx < M;
    unsigned key = 0;
    int val = 0;
    unsigned d = 0;
    if (valid) {
      key = kin[idx];
      val = vin[idx];
      d = (key >> shift) & 255u;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
    __syncthreads();
    unsigned long long mask = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
      const bool on = (d >> bit) & 1u;
      const unsigned long long bb = __ballot(valid && on);
      mask &= on ? bb : ~bb;
    }
    const unsigned rank = (unsigned)__popcll(mask & lt_mask);
    if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(mask);
    __syncthreads();
    if (valid) {
      unsigned pre = run[d] + rank;
      for (int w = 0; w < wave; w++) pre += wcnt[w][d];
      const unsigned at = loc[d] + pre;  // < the block's item count: the counts of all digits sum to it
      sk[at] = key;
      sv[at] = val;
    }
    __syncthreads();
    run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    __syncthreads();
  }
  const int nv = (int)min((long)kSortIPB, M - base);
  for (int i = tid; i < nv; i += kSortT) {
    const unsigned key = sk[i];
    const unsigned pos = gdst[(key >> shift) & 255u] + (unsigned)i;
    kout[pos] = key;
    vout[pos] = sv[i];
  }
}

__global__ void fill_int_k(int* p, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kB) void tile_bounds_k(const unsigned* __restrict__ skeys, long M, unsigned vol,
                                                    int* tile_start) {
  const long i = (long)blockIdx.x * kB + threadIdx.x;
  if (i >= M) return;
  const long t = skeys[i] / vol;
  const long tprev = (i == 0) ? -1 : (long)(skeys[i - 1] / vol);
  for (long tt = tprev + 1; tt <= t; tt++) tile_start[tt] = (int)i;
}

__global__ __launch_bounds__(kB) void tile_chunks_k(const int* __restrict__ tile_start, long ntiles, int dense_min,
                                                    unsigned* chunks, unsigned char* dense) {
  const long t = (long)blockIdx.x * kB + threadIdx.x;
  if (t >= ntiles) return;
  const int cnt = tile_start[t + 1] - tile_start[t];
  const bool d = cnt >= dense_min && cnt > 0;
  dense[t] = d ? 1 : 0;
  chunks[t] = d ? (unsigned)((cnt + kTileNodes - 1) / kTileNodes) : 0u;
}

__global__ __launch_bounds__(kB) void tile_items_k(const unsigned* __restrict__ chunks,
                                                   const unsigned* __restrict__ offs, long ntiles, int* items) {
  const long t = (long)blockIdx.x * kB + threadIdx.x;
  if (t >= ntiles) return;
  const unsigned c = chunks[t], o = offs[t];
  for (unsigned k = 0; k < c; k++) {
    items[2 * (o + k)] = (int)t;
    items[2 * (o + k) + 1] = (int)k;
  }
}

// For each class of the wave's active lanes, one lane adds the class's lane
// count to ctr[class] and every lane of the class gets the old value plus its
// rank among them (lanes in order).  Neighbouring tiles mostly share a class,
// so a wave issues a few atomics instead of one per tile.
__device__ inline unsigned wave_class_add(bool act, int cls, unsigned* ctr) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned pos = 0;
  unsigned long long todo = __ballot(act);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int c = __shfl(cls, leader);
    const unsigned long long same = __ballot(act && cls == c);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(&ctr[c], (unsigned)__popcll(same));
    base = __shfl(base, leader);
    if (act && cls == c) pos = base + (unsigned)__popcll(same & lt_mask);
    todo &= ~same;
  }
  return pos;
}

// Tiles per block of tile_class_k and tile_place_k: the class counts are
// gathered in LDS over this many tiles, so that a block adds each class to the
// global counters once (per-wave global atomics on the 32 counters serialize).
constexpr int kClassTiles = 8 * kB;

// nonempty tiles by node-count class (leading zeros of the count: more nodes,
// smaller class), then placed class by class: most nodes first
__global__ __launch_bounds__(kB) void tile_class_k(const int* __restrict__ tile_start, long ntiles, unsigned* work,
                                                   int* tile_slot) {
  __shared__ unsigned lc[32];
  if (threadIdx.x < 32) lc[threadIdx.x] = 0;
  __syncthreads();
  for (int i = 0; i < kClassTiles / kB; i++) {
    const long t = (long)blockIdx.x * kClassTiles + (long)i * kB + threadIdx.x;
    const int cnt = t < ntiles ? tile_start[t + 1] - tile_start[t] : 0;
    if (t < ntiles) tile_slot[t] = -1;
    (void)wave_class_add(cnt > 0, cnt > 0 ? __clz(cnt) : 0, lc);
  }
  __syncthreads();
  if (threadIdx.x < 32 && lc[threadIdx.x]) atomicAdd(&work[threadIdx.x], lc[threadIdx.x]);
}

// work[0..31] class sizes -> work[32..63] class offsets, work[64] total
__global__ void tile_class_scan_k(unsigned* work) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned acc = 0;
  for (int k = 0; k < 32; k++) {
    work[32 + k] = acc;
    acc += work[k];
  }
  work[64] = acc;
}

// tiles of 2..31 nodes (classes 27..30): flag[t] = 1; tiles of one node (class
// 31): flag[ntiles + t] = 1
__global__ __launch_bounds__(kB) void tile_light_k(const int* __restrict__ tile_start, long ntiles, unsigned* flag) {
  const long t = (long)blockIdx.x * kB + threadIdx.x;
  if (t >= ntiles) return;
  const int cnt = tile_start[t + 1] - tile_start[t];
  flag[t] = (cnt > 1 && cnt < 32) ? 1u : 0u;
  flag[ntiles + t] = cnt == 1 ? 1u : 0u;
}

// Tiles of 32 nodes or more are placed class by class (most nodes first; in a
// class, block by block, a block's tiles in the order of their LDS ranks); the
// tiles of 2..31 nodes follow in tile order (light_off = exclusive scan of
// their flags), so that neighbouring light tiles, whose padded subgrids
// overlap, are taken by neighbouring blocks, and the tiles of one node come
// last, in tile order as well (single_off).
__global__ __launch_bounds__(kB) void tile_place_k(const int* __restrict__ tile_start, long ntiles, unsigned* work,
                                                   const unsigned* __restrict__ light_off,
                                                   const unsigned* __restrict__ single_off, int* order,
                                                   int* tile_slot) {
  constexpr int NI = kClassTiles / kB;
  __shared__ unsigned lc[32], gb[32];
  if (threadIdx.x < 32) lc[threadIdx.x] = 0;
  __syncthreads();
  unsigned lpos[NI];
#pragma unroll
  for (int i = 0; i < NI; i++) {
    const long t = (long)blockIdx.x * kClassTiles + (long)i * kB + threadIdx.x;
    const int cnt = t < ntiles ? tile_start[t + 1] - tile_start[t] : 0;
    lpos[i] = wave_class_add(cnt >= 32, cnt > 0 ? __clz(cnt) : 0, lc);
  }
  __syncthreads();
  if (threadIdx.x < 32 && lc[threadIdx.x]) gb[threadIdx.x] = atomicAdd(&work[32 + threadIdx.x], lc[threadIdx.x]);
  __syncthreads();
  // class 27 starts after the heavy tiles; work[27..30] are the sizes of classes 27..30
  const unsigned light0 = work[32 + 27];
  const unsigned single0 = light0 + work[27] + work[28] + work[29] + work[30];
#pragma unroll
  for (int i = 0; i < NI; i++) {
    const long t = (long)blockIdx.x * kClassTiles + (long)i * kB + threadIdx.x;
    const int cnt = t < ntiles ? tile_start[t + 1] - tile_start[t] : 0;
    if (cnt > 0) {
      const unsigned p = cnt >= 32 ? gb[__clz(cnt)] + lpos[i] : cnt > 1 ? light0 + light_off[t] : single0 + single_off[t];
      order[p] = (int)t;
      tile_slot[t] = (int)p;
    }
  }
}

__global__ void fill_uint_k(unsigned* p, long n, unsigned v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class R>
__global__ void fill_zero_k(Cplx<R>* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = cmk<R>(R(0), R(0));
}

inline unsigned grid1d(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int sort_passes(unsigned long long key_limit) {
  int bits = 0;
  while (bits < 32 && (1ull << bits) < key_limit) bits++;
  return (bits + 7) / 8;
}

static int scan_exclusive(unsigned* d, long n, unsigned* partial, hipStream_t s) {
  const long nch = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(scan_reduce_k, dim3(grid1d(nch * 1, 1)), dim3(kSortT), 0, s, d, n, partial);
  hipLaunchKernelGGL(scan_partials_k, dim3(1), dim3(kSortT), 0, s, partial, nch);
  hipLaunchKernelGGL(scan_apply_k, dim3(grid1d(nch, 1)), dim3(kSortT), 0, s, d, n, partial);
  return kOk;
}

// Bin keys of the M nodes sorted stably (LSD radix, 8-bit digits) with the
// scratch of W: perm_out = node ids in key order, keys_out = the sorted keys.
template <class R>
static int sort_bins(const R* x, long M, const GridGeom& g, const TileLayout& T, SortWork& W, int* perm_out,
                     unsigned* keys_out, hipStream_t s) {
  const unsigned long long key_limit = (unsigned long long)T.ntiles() * (unsigned long long)T.tile_vol();
  if (key_limit > (1ull << 32)) return kUnsupported;
  const int passes = sort_passes(key_limit);
  const int nblk = (int)((M + kSortIPB - 1) / kSortIPB);
  hipLaunchKernelGGL((bin_keys_k<R>), dim3(grid1d(M, kB)), dim3(kB), 0, s, x, M, g, T, W.keys[0], W.vals[0]);
  // the last pass stores straight into the outputs (separate from the scratch)
  int cur = 0;
  for (int ps = 0; ps < passes; ps++) {
    const int shift = 8 * ps;
    const bool last = ps == passes - 1;
    hipLaunchKernelGGL(sort_hist_k, dim3(nblk), dim3(kSortT), 0, s, W.keys[cur], M, shift, W.hist, nblk);
    scan_exclusive(W.hist, 256L * nblk, W.partial, s);
    hipLaunchKernelGGL(sort_scatter_k, dim3(nblk), dim3(kSortT), 0, s, W.keys[cur], W.vals[cur],
                       last ? keys_out : W.keys[1 - cur], last ? perm_out : W.vals[1 - cur], M, shift, W.hist,
                       nblk);
    cur = 1 - cur;
  }
  if (passes == 0) {  // keys below 2: already in order
    if (hipMemcpyAsync(perm_out, W.vals[0], sizeof(int) * M, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return kHipError;
    if (hipMemcpyAsync(keys_out, W.keys[0], sizeof(unsigned) * M, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return kHipError;
  }
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template <class R>
int bin_nodes(const R* x, long M, const GridGeom& g, const TileLayout& T, SortWork& W, int* perm_out,
              unsigned* keys_out, int* tile_start, hipStream_t s) {
  const long ntiles = T.ntiles();
  hipLaunchKernelGGL(fill_int_k, dim3(grid1d(ntiles + 1, 256)), dim3(256), 0, s, tile_start, ntiles + 1, (int)M);
  if (M == 0) return hipGetLastError() == hipSuccess ? kOk : kHipError;
  const int st = sort_bins<R>(x, M, g, T, W, perm_out, keys_out, s);
  if (st != kOk) return st;
  hipLaunchKernelGGL(tile_bounds_k, dim3(grid1d(M, kB)), dim3(kB), 0, s, keys_out, M, (unsigned)T.tile_vol(),
                     tile_start);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

// ---------------------------------------------------------------------------
// binning without allocation: the sort's scratch (two key and two id buffers,
// the digit histograms and their scan partials) lives in one file-scope device
// array per GPU, of kNoAllocMaxNodes nodes at most.  Its address is taken once
// per device with hipGetSymbolAddress (the host-side name is not a device
// address); every call on a GPU shares that GPU's array, so calls that overlap
// on one GPU race.  Every scratch word a call reads was written earlier by the
// same call (the histograms are rebuilt per pass), so nothing needs clearing.
// ---------------------------------------------------------------------------
namespace {
constexpr long kNaBlocks = (kNoAllocMaxNodes + kSortIPB - 1) / kSortIPB;
constexpr long kNaHist = 256L * kNaBlocks;
constexpr long kNaPartial = (kNaHist + kScanChunk - 1) / kScanChunk + 1;
constexpr long kNaWords = 4 * kNoAllocMaxNodes + kNaHist + kNaPartial;
}  // namespace

static __device__ __align__(256) unsigned scratch[kNaWords];

static constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static unsigned* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device (resolved on the first call there).
static unsigned* noalloc_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_na_ptr[dev] = static_cast<unsigned*>(ptr);
  }
  return g_na_ptr[dev];
}

size_t noalloc_scratch_bytes() { return sizeof(unsigned) * (size_t)kNaWords; }

template <class R>
int bin_nodes_noalloc(const R* x, long M, const GridGeom& g, const TileLayout& T, int* perm_out, unsigned* keys_out,
                      hipStream_t s) {
  if (M < 0 || M > kNoAllocMaxNodes) return kUnsupported;
  if (M == 0) return kOk;
  unsigned* base = noalloc_scratch();
  if (!base) return kHipError;
  SortWork W;
  W.keys[0] = base;
  W.keys[1] = base + kNoAllocMaxNodes;
  W.vals[0] = reinterpret_cast<int*>(base + 2 * kNoAllocMaxNodes);
  W.vals[1] = reinterpret_cast<int*>(base + 3 * kNoAllocMaxNodes);
  W.hist = base + 4 * kNoAllocMaxNodes;
  W.partial = W.hist + kNaHist;
  W.cap = kNoAllocMaxNodes;
  W.hist_cap = kNaHist;
  W.partial_cap = kNaPartial;
  return sort_bins<R>(x, M, g, T, W, perm_out, keys_out, s);
}

int build_items(const int* tile_start, long ntiles, int dense_min, unsigned* chunks, unsigned* partial,
                unsigned char* tile_dense, int* items, long items_cap, hipStream_t s) {
  hipLaunchKernelGGL(tile_chunks_k, dim3(grid1d(ntiles, kB)), dim3(kB), 0, s, tile_start, ntiles, dense_min, chunks,
                     tile_dense);
  unsigned last_cnt = 0, last_off = 0;
  if (hipMemcpyAsync(&last_cnt, chunks + ntiles - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess)
    return -1;
  // chunks -> exclusive offsets in place (a copy of the last count was taken above)
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  unsigned* offs = chunks;
  // keep counts: scan a copy held in the second half of the chunk buffer
  if (hipMemcpyAsync(offs + ntiles, chunks, sizeof(unsigned) * ntiles, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return -1;
  scan_exclusive(offs + ntiles, ntiles, partial, s);
  if (hipMemcpyAsync(&last_off, offs + 2 * ntiles - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  const long total = (long)last_off + (long)last_cnt;
  if (total > items_cap) return -1;
  hipLaunchKernelGGL(tile_items_k, dim3(grid1d(ntiles, kB)), dim3(kB), 0, s, chunks, offs + ntiles, ntiles, items);
  if (hipGetLastError() != hipSuccess) return -1;
  return (int)total;
}

int cols_pad(int m, bool is_double, const int* ts) {
  if (m < 2 || m > 8 || (is_double && m > 6)) return 0;
  const int TS = std_tile(m, is_double);
  if (ts[0] != TS || ts[1] != TS || ts[2] != TS) return 0;
  return TS + 2 * m;
}

// Offsets of the nodes of every tile, per axis, as bit masks: the in-tile cell
// index of a sorted key (the layout of bin_keys_k: runs of q = min(TS2, 4)
// cells along axis 2 outermost, then axis 0, axis 1 and the cell within the
// run) is decoded and OR-ed into box[tile] (bits 8 t + l for offset l on axis
// t; tile sizes of at most 8).
// The keys are sorted, so the lanes of a wave that share a tile are
// consecutive: a segmented OR scan over each such run leaves the run's OR in
// its last lane, which makes the one atomic for that tile (instead of one per
// lane on the same word).
__global__ __launch_bounds__(kB) void tile_box_k(const unsigned* __restrict__ keys, long M, TileLayout T,
                                                 unsigned* box) {
  const long i = (long)blockIdx.x * kB + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < M;
  const unsigned vol = (unsigned)(T.ts[0] * T.ts[1] * T.ts[2]);
  const unsigned key = live ? keys[i] : 0u, in = key % vol;
  const unsigned tile = live ? key / vol : 0xFFFFFFFFu;  // no tile id reaches the sentinel
  const unsigned q = T.ts[2] < 4 ? (unsigned)T.ts[2] : 4u;
  const unsigned r = in % q, rest = in / q;  // rest = (l2 / q) * TS0 * TS1 + l0 * TS1 + l1
  const unsigned l1 = rest % (unsigned)T.ts[1], l0 = (rest / (unsigned)T.ts[1]) % (unsigned)T.ts[0];
  const unsigned l2 = (rest / (unsigned)(T.ts[0] * T.ts[1])) * q + r;
  unsigned m = live ? (1u << l0) | (1u << (8 + l1)) | (1u << (16 + l2)) : 0u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned o = (unsigned)__shfl_up((int)m, d);
    const unsigned to = (unsigned)__shfl_up((int)tile, d);
    if (lane >= d && to == tile) m |= o;
  }
  const unsigned tn = (unsigned)__shfl_down((int)tile, 1);
  if (live && (lane == 63 || tn != tile)) atomicOr(box + tile, m);
}

// box_ord[p] = box_tile[order[p]]: the masks by position in the tile order
__global__ __launch_bounds__(kB) void box_order_k(const int* __restrict__ order, int n, const unsigned* __restrict__ box_tile,
                                                  unsigned* box_ord) {
  const int p = blockIdx.x * kB + threadIdx.x;
  if (p < n) box_ord[p] = box_tile[order[p]];
}

int build_tile_boxes(const unsigned* keys, long M, const TileLayout& T, const int* order, int n_order,
                     unsigned* box_tile, unsigned* box_ord, hipStream_t s) {
  if (T.ts[0] > 8 || T.ts[1] > 8 || T.ts[2] > 8) return kUnsupported;
  const long ntiles = T.ntiles();
  hipLaunchKernelGGL(fill_uint_k, dim3(grid1d(ntiles, kB)), dim3(kB), 0, s, box_tile, ntiles, 0u);
  if (M > 0) hipLaunchKernelGGL(tile_box_k, dim3(grid1d(M, kB)), dim3(kB), 0, s, keys, M, T, box_tile);
  if (n_order > 0)
    hipLaunchKernelGGL(box_order_k, dim3(grid1d(n_order, kB)), dim3(kB), 0, s, order, n_order, box_tile, box_ord);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

int build_tile_order(const int* tile_start, long ntiles, unsigned* work, unsigned* light, unsigned* partial,
                     int* order, int* tile_slot, int* n_heavy, int* n_single, hipStream_t s) {
  hipLaunchKernelGGL(fill_uint_k, dim3(1), dim3(128), 0, s, work, 65L, 0u);
  hipLaunchKernelGGL(tile_class_k, dim3(grid1d(ntiles, kClassTiles)), dim3(kB), 0, s, tile_start, ntiles, work,
                     tile_slot);
  hipLaunchKernelGGL(tile_class_scan_k, dim3(1), dim3(64), 0, s, work);
  hipLaunchKernelGGL(tile_light_k, dim3(grid1d(ntiles, kB)), dim3(kB), 0, s, tile_start, ntiles, light);
  scan_exclusive(light, ntiles, partial, s);
  scan_exclusive(light + ntiles, ntiles, partial, s);
  hipLaunchKernelGGL(tile_place_k, dim3(grid1d(ntiles, kClassTiles)), dim3(kB), 0, s, tile_start, ntiles, work,
                     light, light + ntiles, order, tile_slot);
  // work[k] is the size of class k; after the placement work[32 + k] is the end
  // of class k for k <= 26 (32 nodes or more) and work[64] the total
  unsigned w[65] = {0};
  if (hipMemcpyAsync(w, work, 65 * sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  if (n_heavy) *n_heavy = (int)w[32 + 26];
  if (n_single) *n_single = (int)w[31];
  return (int)w[64];
}

template <class R>
int gather(const BArgs<R>& A, int mode, hipStream_t s) {
  if (A.M == 0) return kOk;
  if (A.g.m < kMinPlanM || A.g.m > kMaxPlanM) return kUnsupported;
  switch (mode) {
    case B_F: return b_launch<R, 0, B_F>(A, s);
    case B_GRAD: return b_launch<R, 0, B_GRAD>(A, s);
    case B_FGRAD: return b_launch<R, 0, B_FGRAD>(A, s);
    case B_VJP: return b_launch<R, 0, B_VJP>(A, s);
    default: return kUnsupported;
  }
}

template <class R>
int scatter(const BArgs<R>& A, int mode, hipStream_t s) {
  if (A.M == 0) return kOk;
  if (A.g.m < kMinPlanM || A.g.m > kMaxPlanM) return kUnsupported;
  if (A.slots) {  // column path: writes every grid cell through the fold
    int st = kUnsupported;
    switch (mode) {
      case B_F: st = cols_launch<R, B_F>(A, s); break;
      case B_GRAD: st = cols_launch<R, B_GRAD>(A, s); break;
      case B_FGRAD: st = cols_launch<R, B_FGRAD>(A, s); break;
      default: break;
    }
    if (st != kUnsupported) return st;  // kUnsupported: nothing was launched, take the LDS-atomic path
  }
  // LDS-atomic path: accumulates into the cleared grid
  const long ncell = (long)A.g.n[0] * A.g.n[1] * A.g.n[2];
  const long ntot = A.c2r ? (ncell + 1) / 2 : ncell;  // complex elements of the grid's storage (c2r: real grid)
  hipLaunchKernelGGL((fill_zero_k<R>), dim3(grid1d(ntot, 256)), dim3(256), 0, s, A.grid, (size_t)ntot);
  switch (mode) {
    case B_F: return b_launch<R, 1, B_F>(A, s);
    case B_GRAD: return b_launch<R, 1, B_GRAD>(A, s);
    case B_FGRAD: {  // both contributions add into the same grid
      const int st = b_launch<R, 1, B_F>(A, s);
      return st != kOk ? st : b_launch<R, 1, B_GRAD>(A, s);
    }
    default: return kUnsupported;
  }
}

template <class R>
int fill_zero(Cplx<R>* p, size_t n, hipStream_t s) {
  if (n == 0) return kOk;
  hipLaunchKernelGGL((fill_zero_k<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

template int bin_nodes<float>(const float*, long, const GridGeom&, const TileLayout&, SortWork&, int*, unsigned*,
                              int*, hipStream_t);
template int bin_nodes<double>(const double*, long, const GridGeom&, const TileLayout&, SortWork&, int*, unsigned*,
                               int*, hipStream_t);
template int bin_nodes_noalloc<float>(const float*, long, const GridGeom&, const TileLayout&, int*, unsigned*,
                                      hipStream_t);
template int bin_nodes_noalloc<double>(const double*, long, const GridGeom&, const TileLayout&, int*, unsigned*,
                                       hipStream_t);
template int gather<float>(const BArgs<float>&, int, hipStream_t);
template int gather<double>(const BArgs<double>&, int, hipStream_t);
template int scatter<float>(const BArgs<float>&, int, hipStream_t);
template int scatter<double>(const BArgs<double>&, int, hipStream_t);
template int fill_zero<float>(Cplx<float>*, size_t, hipStream_t);
template int fill_zero<double>(Cplx<double>*, size_t, hipStream_t);

}  // namespace pnfft_amd
