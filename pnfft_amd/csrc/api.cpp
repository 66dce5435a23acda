// api.cpp -- the C interface (include/pnfft.h) on one MI355X.
//
// Plan and node objects mirror plan_s / nodes_s of the reference
// (kernel/ipnfft.h:160-256): the same parameters, the same ownership rules for
// f_hat / x / f / grad_f, and the same stage order in trafo / adj
// (api/api-basic.c:170-378).  The arrays behind the getters may live in host
// memory (the reference's model) or in device memory (set_* with a device
// pointer, or init_nodes_device); host arrays are staged through device
// mirrors on every call, device arrays are used in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_common.hpp"
#include "fft.hpp"
#include "host_math.hpp"
#include "ndft.hpp"
#include "nfft_b.hpp"
#include "window_math.hpp"
#include "../../include/pnfft.h"

// kernel classes of pnfft_kernel_times: the PNFFT_KT_* values of include/pnfft.h

// flag values (api/pnfft.h:302-390)
namespace fl {
constexpr unsigned PRE_PHI_HAT = 1u << 0;
constexpr unsigned MALLOC_F_HAT = 1u << 6;
constexpr unsigned INTERLACED = 1u << 8;
constexpr unsigned TRANSPOSED_F_HAT = 1u << 11;
constexpr unsigned DIFF_IK = 1u << 12;
constexpr unsigned WINDOW_GAUSSIAN = 1u << 13;
constexpr unsigned WINDOW_BESSEL_I0 = 1u << 16;
constexpr unsigned PRE_CONST_PSI = 1u << 2;
constexpr unsigned PRE_LIN_PSI = 1u << 3;
constexpr unsigned PRE_QUAD_PSI = 1u << 4;
constexpr unsigned PRE_CUB_PSI = 1u << 5;
constexpr unsigned USE_FK_GAUSSIAN_T = 1u << 17;
constexpr unsigned REAL_F = 1u << 4;
constexpr unsigned PRE_GRAD_PSI = 1u << 2;
constexpr unsigned PRE_HESSIAN_PSI = 1u << 3;
constexpr unsigned PRE_PSI = 1u << 1;
constexpr unsigned MALLOC_X = 1u << 0;
constexpr unsigned MALLOC_F = 1u << 1;
constexpr unsigned MALLOC_GRAD_F = 1u << 2;
constexpr unsigned MALLOC_HESSIAN_F = 1u << 3;
constexpr unsigned COMPUTE_F = 1u << 0;
constexpr unsigned COMPUTE_GRAD_F = 1u << 1;
constexpr unsigned COMPUTE_HESSIAN_F = 1u << 2;
constexpr unsigned COMPUTE_DIRECT = 1u << 3;
constexpr unsigned COMPUTE_ACCUMULATED = 1u << 4;
constexpr unsigned OMIT_DECONV = 1u << 5;
constexpr unsigned OMIT_FFT = 1u << 6;
constexpr unsigned OMIT_CONV = 1u << 7;
}  // namespace fl

enum TimerSlot {
  T_ITER = 0, T_WHOLE = 1, T_LOOP_B = 2, T_SORT = 3, T_GCELLS = 4, T_B = 5, T_F = 6, T_D = 7, T_SHIFT_IN = 8,
  T_SHIFT_OUT = 9, T_LEN = 10
};

namespace pnfft_amd {

// ---------------------------------------------------------------------------
// errors and memory registry
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

// Each API call that can fail starts with a clean state, so last_error names
// the failure of the most recent such call (or is NULL after a success).
static void clear_error() { g_last_error.clear(); }

static void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  fprintf(stderr, "!!! Error in PNFFT: %s !!!\n", buf);
}

enum MemKind : int { MEM_PINNED = 1, MEM_DEVICE = 2, MEM_HEAP = 3 };
static std::mutex g_mem_mu;
static std::unordered_map<void*, int>& mem_registry() {
  static std::unordered_map<void*, int> m;
  return m;
}

static void* host_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    p = aligned_alloc(64, (bytes + 63) / 64 * 64);
    if (!p) return nullptr;
    std::lock_guard<std::mutex> g(g_mem_mu);
    mem_registry()[p] = MEM_HEAP;
    return p;
  }
  std::lock_guard<std::mutex> g(g_mem_mu);
  mem_registry()[p] = MEM_PINNED;
  return p;
}

static void* device_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of %zu bytes failed", bytes);
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_mem_mu);
  mem_registry()[p] = MEM_DEVICE;
  return p;
}

static void mem_free(void* p) {
  if (!p) return;
  int kind = 0;
  {
    std::lock_guard<std::mutex> g(g_mem_mu);
    auto it = mem_registry().find(p);
    if (it != mem_registry().end()) {
      kind = it->second;
      mem_registry().erase(it);
    }
  }
  if (kind == MEM_PINNED) (void)hipHostFree(p);
  else if (kind == MEM_DEVICE) (void)hipFree(p);
  else if (kind == MEM_HEAP) free(p);
  else set_error("free of a pointer that was not allocated by the library");
}

// internal device buffers (not registered: freed directly)
template <class T>
static T* dalloc(size_t count) {
  if (count == 0) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return static_cast<T*>(p);
}
template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

// ---------------------------------------------------------------------------
// plan and nodes
// ---------------------------------------------------------------------------
// Event slots: 0/1 binning, 2/3 FFT (8, 9 after its first and second pass),
// 4/5 B (10 and 13 between its kernels).
struct EventTimer {
  hipEvent_t ev[16] = {};
  bool ok = false;
  void init() {
    ok = true;
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) ok = false;
  }
  void release() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void rec(int i, hipStream_t s) {
    if (ok) (void)hipEventRecord(ev[i], s);
  }
  hipEvent_t mark(int i) const { return ok ? ev[i] : nullptr; }
  double ms(int a, int b) {
    float t = 0;
    if (!ok || hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) {
      (void)hipGetLastError();
      return 0.0;
    }
    return t;
  }
};

// the opaque handle types of include/pnfft.h for each precision (defined with the C interface)
template <class R>
struct Handles;

template <class R>
struct PlanT {
  int d = 3;
  long N[3] = {0, 0, 0}, n[3] = {0, 0, 0}, no[3] = {0, 0, 0};
  int m = 6;
  double x_max[3] = {0.5, 0.5, 0.5};
  double sigma[3] = {2, 2, 2};
  double b[3] = {0, 0, 0};
  unsigned pnfft_flags = 0, pfft_flags = 0;
  int win = WIN_KB;
  long N_total = 0, n_total = 0;
  bool c2r = false;     // real plan (PNX(init_guru_c2r)): f_hat is the half spectrum k2 in [-N2/2, 0]
  long fhat_elems = 0;  // complex entries of f_hat: N_total, or N0*N1*(N2/2+1) for c2r

  void* f_hat = nullptr;  // user-visible pointer (host or device)
  bool own_f_hat = false;

  hipStream_t stream = nullptr;
  Fft3d<R> fft;
  Cplx<R>* d_fhat = nullptr;
  Cplx<R>* d_grid = nullptr;
  Cplx<R>* d_buf1 = nullptr;
  Cplx<R>* d_buf2 = nullptr;
  double* d_tab = nullptr;
  std::vector<double> h_tab;
  double* d_tab_scale = nullptr;  // 2^e per axis and frequency: D of PNFFT_OMIT_DECONV (window scale only)
  // complex per-axis factors of interlaced / ik-differentiated passes (kCtabSlots
  // tables of N0 + N1 + N2 (re, im) pairs, filled per call in pinned host memory)
  double* d_ctab = nullptr;
  double* h_ctab = nullptr;
  bool gauss_t = false;   // PNFFT_WINDOW_GAUSSIAN_T: the truncated Gaussian's D (matrix_D.c:37-64)
  // windows the B kernels take from per-call tables (window_ext.hpp): the
  // Bessel-I0 window and the interpolated windows of PNFFT_PRE_*_PSI
  bool tab_only = false;
  IntpolTabs<R> ip{};
  R* d_ip = nullptr;      // interpolation tables, 9 of them ([derivative][axis])
  GridGeom geom{};
  TileLayout tiles{};
  EventTimer evt;
  double timer_trafo[T_LEN] = {0};
  double timer_adj[T_LEN] = {0};
  double ktime[PNFFT_KT_COUNT] = {0};  // per-kernel-class device ms (pnfft_kernel_times)
  long kcount[PNFFT_KT_COUNT] = {0};
};

template <class R>
struct NodesT {
  long M = 0;
  R* x = nullptr;
  R* f = nullptr;
  R* grad_f = nullptr;
  R* hessian_f = nullptr;
  bool device_arrays = false;  // created by init_nodes_device
  unsigned precompute_flags = 0;

  // device mirrors of host arrays
  R* d_x = nullptr;
  Cplx<R>* d_f = nullptr;
  Cplx<R>* d_grad = nullptr;
  Cplx<R>* d_hess = nullptr;
  R* d_gx = nullptr;
  // scratch of the ik-differentiated transforms (one component of node values)
  // and of the Hessian gather (window table, stencil starts)
  Cplx<R>* d_tmp = nullptr;
  R* d_hess_tab = nullptr;
  int* d_hess_st = nullptr;
  size_t tmp_cap = 0, hess_tab_cap = 0, hess_st_cap = 0;

  // binning
  int* d_perm = nullptr;
  unsigned* d_keys = nullptr;
  int* d_tile_start = nullptr;
  long tile_cap = 0;
  unsigned* d_chunks = nullptr;      // 2 * ntiles: per-tile chunk counts and their offsets
  unsigned* d_tpartial = nullptr;    // scan partials of the tile counts
  unsigned char* d_dense = nullptr;  // gather: per-tile LDS-path flag
  int* d_items = nullptr;            // gather: (tile, chunk) pairs
  int n_items = 0;
  unsigned char* d_dense_s = nullptr;  // scatter: every nonempty tile takes the LDS path
  int* d_items_s = nullptr;
  int n_items_s = 0;
  long items_cap = 0;
  // column scatter: nonempty tiles by decreasing node count, tile -> slot map, slots
  int* d_torder = nullptr;
  int* d_tile_slot = nullptr;
  unsigned* d_twork = nullptr;
  Cplx<R>* d_slots = nullptr;
  size_t slots_cap = 0;  // complex elements
  int n_torder = 0;
  int n_heavy = 0;       // torder[0, n_heavy): tiles with 32 nodes or more
  int n_single = 0;      // the last n_single tiles of torder: one node each
  // offsets of each tile's nodes per axis (build_tile_boxes), by tile id and by position in torder
  unsigned* d_box_tile = nullptr;
  unsigned* d_box_ord = nullptr;
  bool boxes = false;
  int* d_stale = nullptr;            // [M] sorted positions of nodes outside their tile
  unsigned* d_stale_count = nullptr;
  // window tables of precompute_psi (sorted order, see precompute_tables)
  R* d_pre_psi = nullptr;
  R* d_pre_dpsi = nullptr;
  R* d_pre_x = nullptr;
  size_t pre_psi_cap = 0, pre_dpsi_cap = 0, pre_x_cap = 0;
  bool pre_valid = false, pre_grad = false;
  GridGeom pre_geom{};
  SortWork sw;
  bool bins_valid = false;
  GridGeom bin_geom{};
  TileLayout bin_tiles{};
};

// 1/phi_hat and phi_hat of axis t (PNX(inv_phi_hat), PNX(phi_hat), kernel/matrix_D.c:191-225)
template <class R>
static double plan_phi_hat(const PlanT<R>* P, int t, long k) {
  if (P->gauss_t) return host::phi_hat_gauss_t(k, P->n[t], P->b[t], P->m);
  return host::phi_hat(P->win, k, P->n[t], P->b[t], P->m);
}
template <class R>
static double plan_inv_phi_hat(const PlanT<R>* P, int t, long k) {
  if (P->gauss_t) return 1.0 / host::phi_hat_gauss_t(k, P->n[t], P->b[t], P->m);
  return host::inv_phi_hat(P->win, k, P->n[t], P->b[t], P->m);
}

// Interpolation order of kernel/ndft-parallel.c:997-1006 (on the plan flags
// after the implications of api/api-guru.c:152-155, so LIN implies CONST).
static int intpol_order_of(unsigned flags) {
  if (flags & fl::PRE_CONST_PSI) return 0;
  if (flags & fl::PRE_LIN_PSI) return 1;
  if (flags & fl::PRE_QUAD_PSI) return 2;
  if (flags & fl::PRE_CUB_PSI) return 3;
  return -1;
}

// Interpolation tables of init_intpol_table_psi (kernel/ndft-parallel.c:318-352):
// K = ceil(31/(2m+1)) 2048 nodes per unit interval (:1095); derivative tables
// only without ik-differentiation (:1106-1121).  Entries shared by neighbouring
// k are evaluated once, as the reference does.
template <class R>
static bool intpol_tables(PlanT<R>* P, int order) {
  const int m = P->m, cutoff = 2 * m + 1, w = order + 1;
  const long K = (long)std::ceil(31.0 / cutoff) * 2048;
  const size_t per = (size_t)K * cutoff * w;
  const int nderiv = (P->pnfft_flags & fl::DIFF_IK) ? 1 : 3;
  std::vector<R> h(per * 9, R(0));
  for (int d = 0; d < nderiv; d++)
    for (int t = 0; t < 3; t++) {
      R* tab = h.data() + (size_t)(d * 3 + t) * per;
      size_t ind = 0;
      for (long k = 0; k < K; k++)
        for (int c = 0; c < cutoff; c++)
          for (int i = -order / 2; i <= (order + 1) / 2; i++, ind++) {
            if (k > 0 && i < (order + 1) / 2) {
              tab[ind] = tab[ind - (size_t)cutoff * w + 1];
              continue;
            }
            const double x = ((double)m + (double)(k + i) / (double)K - c) / (double)P->n[t];
            const double v = d == 0   ? host::psi(P->win, P->n[t], P->b[t], m, x)
                             : d == 1 ? host::dpsi(P->win, P->n[t], P->b[t], m, x)
                                      : host::ddpsi(P->win, P->n[t], P->b[t], m, x);
            tab[ind] = (R)v;
          }
    }
  P->d_ip = dalloc<R>(per * 9);
  if (!P->d_ip || hipMemcpy(P->d_ip, h.data(), sizeof(R) * per * 9, hipMemcpyHostToDevice) != hipSuccess)
    return false;
  P->ip.order = order;
  P->ip.K = K;
  for (int d = 0; d < 3; d++)
    for (int t = 0; t < 3; t++) P->ip.tab[d][t] = P->d_ip + (size_t)(d * 3 + t) * per;
  return true;
}

template <class R>
static void tables_compute(PlanT<R>* P) {
  // window scale 2^-e per axis, e = exponent of psi(0) (the window's peak); the
  // D factors carry 2^e, so B F D is unchanged (kernel/matrix_D.c:227-280)
  int e[3];
  for (int t = 0; t < 3; t++) {
    const double peak = std::fabs(host::psi(P->win, P->n[t], P->b[t], P->m, 0.0));
    e[t] = (std::isfinite(peak) && peak > 0.0) ? std::ilogb(peak) : 0;
    P->geom.ws[t] = std::ldexp(1.0, -e[t]);
  }
  P->h_tab.resize(P->N[0] + P->N[1] + P->N[2]);
  long off = 0;
  for (int t = 0; t < 3; t++)
    for (long i = 0; i < P->N[t]; i++)
      P->h_tab[off++] = std::ldexp(plan_inv_phi_hat(P, t, i - P->N[t] / 2), e[t]);
  if (P->d_tab)
    (void)hipMemcpy(P->d_tab, P->h_tab.data(), sizeof(double) * P->h_tab.size(), hipMemcpyHostToDevice);
  std::vector<double> sc(P->h_tab.size());
  off = 0;
  for (int t = 0; t < 3; t++)
    for (long i = 0; i < P->N[t]; i++) sc[off++] = std::ldexp(1.0, e[t]);
  if (P->d_tab_scale)
    (void)hipMemcpy(P->d_tab_scale, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice);
}

// Slots of the complex factor tables: per pass (interlaced plans have two), the
// transform itself (0), the three gradient components (1..3) and the six
// Hessian components (4..9, order xx, xy, xz, yy, yz, zz).
constexpr int kCtabComps = 10;
constexpr int kCtabSlots = 2 * kCtabComps;
static const int kHessA[6] = {0, 0, 0, 1, 1, 2}, kHessB[6] = {0, 1, 2, 1, 2, 2};

// Complex per-axis factors of one pass (FP64, f_hat's index order, N0 + N1 + N2
// (re, im) pairs), their product being the pass's factor at (k0, k1, k2):
//  * D (matrix_D.c:227-280) with the window scale 2^e of the plan, or the scale
//    alone under PNFFT_OMIT_DECONV;
//  * interlaced plans: 1/2 (assign.c:688, spread :481) and, for the shifted pass,
//    the phase e^{+pi i k_t/n_t} (trafo) / e^{-pi i k_t/n_t} (adj) of
//    convolution_due_to_interlacing (matrix_D.c:282-317);
//  * ik-differentiation (ndft-parallel.c:3012-3087): -2 pi i k_t (trafo) or
//    +2 pi i k_t (adj) on the derivative's axes (comp 1..3: axis comp-1; comp
//    4..9: the Hessian pair, -4 pi^2 k_a k_b when both are taken).
template <class R>
static void ctab_fill(const PlanT<R>* P, bool shifted, bool adjoint, int comp, bool omit_d, double* out) {
  const bool il = (P->pnfft_flags & fl::INTERLACED) != 0;
  int da = -1, db = -1;
  if (comp >= 1 && comp <= 3) da = comp - 1;
  if (comp >= 4) {
    da = kHessA[comp - 4];
    db = kHessB[comp - 4];
  }
  const double pi = 3.14159265358979323846;
  long off = 0;
  for (int t = 0; t < 3; t++) {
    const double e = std::ldexp(1.0, std::ilogb(1.0 / P->geom.ws[t]));
    for (long i = 0; i < P->N[t]; i++, off++) {
      const double k = (double)(i - P->N[t] / 2);
      double re = omit_d ? e : P->h_tab[off], im = 0.0;
      if (il && t == 0) re *= 0.5;
      auto mul = [&](double cr, double ci) {
        const double r = re * cr - im * ci;
        im = re * ci + im * cr;
        re = r;
      };
      if (shifted) {
        const double ph = (adjoint ? -pi : pi) * k / (double)P->n[t];
        mul(std::cos(ph), std::sin(ph));
      }
      const double dk = (adjoint ? 2.0 : -2.0) * pi * k;
      if (t == da) mul(0.0, dk);
      if (t == db) mul(0.0, dk);
      out[2 * off] = re;
      out[2 * off + 1] = im;
    }
  }
}

// Tile layout for binning: std_tile cells per axis (8 or 4, fewer on shorter axes).
static TileLayout choose_tiles(const long* n, int m, bool is_double) {
  TileLayout T{};
  for (int t = 0; t < 3; t++) {
    int ts = std_tile(m, is_double);
    while (ts > 1 && n[t] % ts) ts >>= 1;
    T.ts[t] = ts;
    T.ntile[t] = (int)((n[t] + ts - 1) / ts);
  }
  return T;
}

template <class R>
static void plan_destroy(PlanT<R>* P) {
  if (!P) return;
  if (P->stream) (void)hipStreamSynchronize(P->stream);
  P->fft.release();
  dfree(P->d_fhat);
  dfree(P->d_grid);
  dfree(P->d_buf1);
  dfree(P->d_buf2);
  dfree(P->d_tab);
  dfree(P->d_tab_scale);
  dfree(P->d_ctab);
  if (P->h_ctab) (void)hipHostFree(P->h_ctab);
  dfree(P->d_ip);
  P->evt.release();
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete static_cast<typename Handles<R>::plan*>(P);
}

// api/api-guru.c:110-164 + kernel/ndft-parallel.c:869-1069
template <class R>
static PlanT<R>* plan_create(int d, const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m,
                             unsigned pnfft_flags, unsigned pfft_flags, bool c2r) {
  clear_error();
  if (d != 3) {
    set_error("d != 3 not yet implemented");
    return nullptr;
  }
  if (c2r && (N[2] % 2)) {
    set_error("c2r plans need an even N[2] (half spectrum N[2]/2+1)");
    return nullptr;
  }
  const unsigned unsupported = fl::TRANSPOSED_F_HAT;
  if (pnfft_flags & unsupported) {
    set_error("plan flags 0x%x (transposed f_hat) are not supported by this build", pnfft_flags & unsupported);
    return nullptr;
  }
  // the B kernels are compiled for stencils of 2m+1 = 5..17 points
  if (m < kMinPlanM || m > kMaxPlanM) {
    set_error("window cutoff m=%d outside [%d, %d] (the cutoffs this build's kernels are compiled for)", m,
              kMinPlanM, kMaxPlanM);
    return nullptr;
  }
  // FP64 c2r plans with m >= 7 stalled a GPU run of an earlier kernel build and
  // have not been run since; they are refused until that is understood
  if (c2r && sizeof(R) == 8 && m >= 7) {
    set_error("double-precision c2r plans with m=%d >= 7 are not supported by this build", m);
    return nullptr;
  }
  for (int t = 0; t < 3; t++) {
    if (N[t] < 1 || n[t] < N[t]) {
      set_error("need 1 <= N[%d]=%td <= n[%d]=%td", t, N[t], t, n[t]);
      return nullptr;
    }
    if (!fft_length_supported((int)n[t], sizeof(R) == 8)) {
      set_error("oversampled length n[%d]=%td must be a power of two <= %d", t, n[t], sizeof(R) == 8 ? 4096 : 8192);
      return nullptr;
    }
    if (n[t] < 2 * m + 1) {
      set_error("oversampled length n[%d]=%td is shorter than the stencil 2m+1=%d", t, n[t], 2 * m + 1);
      return nullptr;
    }
  }
  // precomputation implications of api/api-guru.c:152-155
  if (pnfft_flags & fl::PRE_HESSIAN_PSI) pnfft_flags |= fl::PRE_GRAD_PSI;
  if (pnfft_flags & fl::PRE_GRAD_PSI) pnfft_flags |= fl::PRE_PSI;

  PlanT<R>* P = new typename Handles<R>::plan();
  P->m = m;
  P->pnfft_flags = pnfft_flags;
  P->pfft_flags = pfft_flags;
  P->win = host::window_from_flags(pnfft_flags);
  P->gauss_t = (pnfft_flags & fl::WINDOW_GAUSSIAN) && (pnfft_flags & fl::USE_FK_GAUSSIAN_T);
  const int ip_order = intpol_order_of(pnfft_flags);
  P->tab_only = P->win == WIN_BESSEL_I0 || ip_order >= 0;
  P->N_total = P->n_total = 1;
  for (int t = 0; t < 3; t++) {
    P->N[t] = N[t];
    P->n[t] = n[t];
    P->x_max[t] = (double)x_max[t];
    // fft_output_size, api/api-guru.c:195-218
    const long c = (long)std::lrint(std::floor((double)n[t] * (double)x_max[t]));
    P->no[t] = std::min<long>(n[t], 2 * (c + m + 2));
    P->sigma[t] = (double)n[t] / (double)N[t];
    P->b[t] = host::window_b(P->win, m, P->sigma[t]);
    P->N_total *= N[t];
    P->n_total *= n[t];
  }
  P->c2r = c2r;
  P->fhat_elems = c2r ? (long)N[0] * N[1] * (N[2] / 2 + 1) : P->N_total;
  P->geom.m = m;
  P->geom.win = P->win;
  for (int t = 0; t < 3; t++) {
    P->geom.n[t] = (int)n[t];
    P->geom.b[t] = P->b[t];
  }
  P->tiles = choose_tiles(P->n, m, sizeof(R) == 8);

  bool ok = hipStreamCreate(&P->stream) == hipSuccess;
  int Ni[3] = {(int)N[0], (int)N[1], (int)N[2]}, ni[3] = {(int)n[0], (int)n[1], (int)n[2]};
  if (ok) ok = P->fft.init(Ni, ni, c2r) == kOk;
  if (ok) {
    P->d_fhat = dalloc<Cplx<R>>(P->N_total);
    P->d_grid = dalloc<Cplx<R>>(P->fft.grid_elems());  // c2r plans: n_total reals
    P->d_buf1 = dalloc<Cplx<R>>(P->fft.buf1_elems());
    P->d_buf2 = dalloc<Cplx<R>>(P->fft.buf2_elems());
    P->d_tab = dalloc<double>(N[0] + N[1] + N[2]);
    P->d_tab_scale = dalloc<double>(N[0] + N[1] + N[2]);
    const size_t ct = (size_t)kCtabSlots * 2 * (N[0] + N[1] + N[2]);
    P->d_ctab = dalloc<double>(ct);
    if (hipHostMalloc((void**)&P->h_ctab, sizeof(double) * ct, 0) != hipSuccess) P->h_ctab = nullptr;
    ok = P->d_fhat && P->d_grid && P->d_buf1 && P->d_buf2 && P->d_tab && P->d_tab_scale && P->d_ctab && P->h_ctab;
  }
  if (!ok) {
    (void)hipGetLastError();
    set_error("device setup failed (memory or HIP runtime)");
    plan_destroy(P);
    return nullptr;
  }
  P->evt.init();
  tables_compute(P);
  if (ip_order >= 0 && !intpol_tables(P, ip_order)) {
    (void)hipGetLastError();
    set_error("device setup of the interpolation tables failed");
    plan_destroy(P);
    return nullptr;
  }
  if (pnfft_flags & fl::MALLOC_F_HAT) {
    P->f_hat = host_alloc(sizeof(Cplx<R>) * P->fhat_elems);
    P->own_f_hat = true;
  }
  return P;
}

template <class R>
static void nodes_release_device(NodesT<R>* D) {
  dfree(D->d_x);
  dfree(D->d_f);
  dfree(D->d_grad);
  dfree(D->d_hess);
  dfree(D->d_gx);
  dfree(D->d_tmp);
  dfree(D->d_hess_tab);
  dfree(D->d_hess_st);
  D->tmp_cap = D->hess_tab_cap = D->hess_st_cap = 0;
  dfree(D->d_perm);
  dfree(D->d_keys);
  dfree(D->d_tile_start);
  dfree(D->d_chunks);
  dfree(D->d_tpartial);
  dfree(D->d_dense);
  dfree(D->d_items);
  dfree(D->d_dense_s);
  dfree(D->d_items_s);
  D->items_cap = 0;
  D->n_items = D->n_items_s = 0;
  dfree(D->d_torder);
  dfree(D->d_tile_slot);
  dfree(D->d_twork);
  dfree(D->d_box_tile);
  dfree(D->d_box_ord);
  D->boxes = false;
  dfree(D->d_slots);
  D->slots_cap = 0;
  D->n_torder = 0;
  dfree(D->d_stale);
  dfree(D->d_stale_count);
  dfree(D->d_pre_psi);
  dfree(D->d_pre_dpsi);
  dfree(D->d_pre_x);
  D->pre_psi_cap = D->pre_dpsi_cap = D->pre_x_cap = 0;
  D->pre_valid = false;
  dfree(D->sw.keys[0]);
  dfree(D->sw.keys[1]);
  dfree(D->sw.vals[0]);
  dfree(D->sw.vals[1]);
  dfree(D->sw.hist);
  dfree(D->sw.partial);
  D->sw.cap = D->sw.hist_cap = D->sw.partial_cap = 0;
  D->tile_cap = 0;
  D->bins_valid = false;
}

template <class R>
static void nodes_destroy(NodesT<R>* D) {
  nodes_release_device(D);
  delete static_cast<typename Handles<R>::nodes*>(D);
}

template <class R>
static NodesT<R>* nodes_create(long M, unsigned malloc_flags, bool device) {
  NodesT<R>* D = new typename Handles<R>::nodes();
  D->M = M;
  D->device_arrays = device;
  auto alloc = [&](size_t bytes) -> R* {
    if (M <= 0) return nullptr;  // kernel/ndft-parallel.c:1491
    return static_cast<R*>(device ? device_alloc(bytes) : host_alloc(bytes));
  };
  if (malloc_flags & fl::MALLOC_X) D->x = alloc(sizeof(R) * 3 * M);
  if (malloc_flags & fl::MALLOC_F) D->f = alloc(sizeof(R) * 2 * M);
  if (malloc_flags & fl::MALLOC_GRAD_F) D->grad_f = alloc(sizeof(R) * 6 * M);
  if (malloc_flags & fl::MALLOC_HESSIAN_F) D->hessian_f = alloc(sizeof(R) * 12 * M);
  return D;
}

// Resolve an array for the kernels: device pointers are used in place, host
// arrays get a device mirror (allocated on demand, uploaded when `upload`).
// A mirror always holds `count` elements of T (the complex size for node
// values, whichever plan type uses the node set); `copy_bytes` (default: all)
// is what is transferred, e.g. the M reals of a c2r plan's f.
template <class T>
static T* stage_in(void* user, T*& mirror, size_t count, bool upload, hipStream_t s, bool* is_host,
                   size_t copy_bytes = 0) {
  *is_host = false;
  if (!user) return nullptr;
  if (is_device_ptr(user)) return static_cast<T*>(user);
  *is_host = true;
  if (!mirror) mirror = dalloc<T>(count);
  if (!mirror) return nullptr;
  if (upload)
    (void)hipMemcpyAsync(mirror, user, copy_bytes ? copy_bytes : sizeof(T) * count, hipMemcpyHostToDevice, s);
  return mirror;
}

template <class R>
static bool ensure_sort_capacity(NodesT<R>* D, const TileLayout& T) {
  const long M = D->M;
  if (D->sw.cap < M) {
    dfree(D->sw.keys[0]);
    dfree(D->sw.keys[1]);
    dfree(D->sw.vals[0]);
    dfree(D->sw.vals[1]);
    dfree(D->d_perm);
    dfree(D->d_keys);
    dfree(D->d_stale);
    dfree(D->d_stale_count);
    D->sw.keys[0] = dalloc<unsigned>(M);
    D->sw.keys[1] = dalloc<unsigned>(M);
    D->sw.vals[0] = dalloc<int>(M);
    D->sw.vals[1] = dalloc<int>(M);
    D->d_perm = dalloc<int>(M);
    D->d_keys = dalloc<unsigned>(M);
    D->d_stale = dalloc<int>(M);
    D->d_stale_count = dalloc<unsigned>(1);
    if (!D->sw.keys[0] || !D->sw.keys[1] || !D->sw.vals[0] || !D->sw.vals[1] || !D->d_perm || !D->d_keys ||
        !D->d_stale || !D->d_stale_count)
      return false;
    D->sw.cap = M;
  }
  const long nblk = (M + 4095) / 4096;
  const long hist = 256 * nblk;
  if (D->sw.hist_cap < hist) {
    dfree(D->sw.hist);
    D->sw.hist = dalloc<unsigned>(hist);
    if (!D->sw.hist) return false;
    D->sw.hist_cap = hist;
  }
  const long parts = (hist + 4095) / 4096;
  if (D->sw.partial_cap < parts) {
    dfree(D->sw.partial);
    D->sw.partial = dalloc<unsigned>(parts);
    if (!D->sw.partial) return false;
    D->sw.partial_cap = parts;
  }
  if (D->tile_cap < T.ntiles() + 1) {
    dfree(D->d_tile_start);
    dfree(D->d_chunks);
    dfree(D->d_tpartial);
    dfree(D->d_dense);
    dfree(D->d_dense_s);
    dfree(D->d_torder);
    dfree(D->d_tile_slot);
    dfree(D->d_twork);
    dfree(D->d_box_tile);
    dfree(D->d_box_ord);
    D->d_tile_start = dalloc<int>(T.ntiles() + 1);
    D->d_chunks = dalloc<unsigned>(2 * T.ntiles());
    D->d_tpartial = dalloc<unsigned>((T.ntiles() + 4095) / 4096 + 1);
    D->d_dense = dalloc<unsigned char>(T.ntiles());
    D->d_dense_s = dalloc<unsigned char>(T.ntiles());
    D->d_torder = dalloc<int>(T.ntiles());
    D->d_tile_slot = dalloc<int>(T.ntiles());
    D->d_twork = dalloc<unsigned>(65);
    D->d_box_tile = dalloc<unsigned>(T.ntiles());
    D->d_box_ord = dalloc<unsigned>(T.ntiles());
    if (!D->d_tile_start || !D->d_chunks || !D->d_tpartial || !D->d_dense || !D->d_dense_s || !D->d_torder ||
        !D->d_tile_slot || !D->d_twork || !D->d_box_tile || !D->d_box_ord)
      return false;
    D->tile_cap = T.ntiles() + 1;
  }
  const long icap = T.ntiles() + (M + kTileNodes - 1) / kTileNodes + 1;
  if (D->items_cap < icap) {
    dfree(D->d_items);
    dfree(D->d_items_s);
    D->d_items = dalloc<int>(2 * icap);
    D->d_items_s = dalloc<int>(2 * icap);
    if (!D->d_items || !D->d_items_s) return false;
    D->items_cap = icap;
  }
  return true;
}

// Scatter path: the column scatter (default) or, for comparison, the
// LDS-atomic kernels (PNFFT_AMD_SCATTER=atomic).  PNFFT_AMD_MFMA=0 keeps the
// vector kernels where the matrix-core gather and scatter would apply.
static bool cols_enabled() {
  static const bool v = [] {
    const char* e = getenv("PNFFT_AMD_SCATTER");
    return !(e && strcmp(e, "atomic") == 0);
  }();
  return v;
}
// PNFFT_AMD_TABLES=0: precompute_psi bins the nodes but tabulates no windows
// (the B kernels evaluate them), for comparison; read on every precompute call
static bool tables_enabled() {
  const char* e = getenv("PNFFT_AMD_TABLES");
  return !(e && strcmp(e, "0") == 0);
}
static bool mfma_disabled() {
  static const bool v = [] {
    const char* e = getenv("PNFFT_AMD_MFMA");
    return e && strcmp(e, "0") == 0;
  }();
  return v;
}

// Minimum node count for a tile to take the LDS path (tunable by PNFFT_AMD_DENSE_MIN).
static int dense_min() {
  static int v = [] {
    const char* e = getenv("PNFFT_AMD_DENSE_MIN");
    return e ? std::max(1, atoi(e)) : 1;
  }();
  return v;
}

template <class R>
static void set_tiles(BArgs<R>& A, const NodesT<R>* D, bool for_scatter = false) {
  A.tile_start = D->d_tile_start;
  A.keys = D->d_keys;
  A.tile_dense = for_scatter ? D->d_dense_s : D->d_dense;
  A.items = for_scatter ? D->d_items_s : D->d_items;
  A.n_items = for_scatter ? D->n_items_s : D->n_items;
  // the scatter's list holds every nonempty tile; the gather's when dense_min is 1
  A.all_dense = (for_scatter || dense_min() == 1) ? 1 : 0;
  A.no_mfma = mfma_disabled() ? 1 : 0;
  if (D->n_torder > 0) {
    A.torder = D->d_torder;
    A.n_torder = D->n_torder;
    A.n_heavy = D->n_heavy;
    A.n_single = D->n_single;
    A.stale_count = D->d_stale_count;
    A.stale_list = D->d_stale;
  }
  if (for_scatter && D->n_torder > 0) {
    A.slots = D->d_slots;
    A.tile_slot = D->d_tile_slot;
  }
  if (D->n_torder > 0 && D->boxes) {
    A.box_tile = D->d_box_tile;
    A.box_ord = D->d_box_ord;
  }
  for (int t = 0; t < 3; t++) {
    A.ts[t] = D->bin_tiles.ts[t];
    A.ntile[t] = D->bin_tiles.ntile[t];
  }
}

static bool same_geom(const GridGeom& a, const GridGeom& b) {
  return a.n[0] == b.n[0] && a.n[1] == b.n[1] && a.n[2] == b.n[2] && a.m == b.m && a.shift == b.shift;
}

// Bin the nodes of D for plan P (on P's stream) in geometry g (the plan's, or
// the half-cell shifted one of an interlaced pass).
template <class R>
static int bin(PlanT<R>* P, NodesT<R>* D, const R* dx, const GridGeom& g) {
  if (!ensure_sort_capacity(D, P->tiles)) {
    set_error("device allocation for node binning failed");
    return kHipError;
  }
  int st = bin_nodes<R>(dx, D->M, g, P->tiles, D->sw, D->d_perm, D->d_keys, D->d_tile_start, P->stream);
  if (st != kOk) {
    set_error("node binning failed (%d)", st);
    return st;
  }
  D->n_items = D->n_items_s = 0;
  if (P->m >= 2 && P->m <= 8) {
    const int ni = build_items(D->d_tile_start, P->tiles.ntiles(), dense_min(), D->d_chunks, D->d_tpartial,
                               D->d_dense, D->d_items, D->items_cap, P->stream);
    const int ns = build_items(D->d_tile_start, P->tiles.ntiles(), 1, D->d_chunks, D->d_tpartial, D->d_dense_s,
                               D->d_items_s, D->items_cap, P->stream);
    if (ni < 0 || ns < 0) {
      set_error("building the tile work list failed");
      return kHipError;
    }
    D->n_items = ni;
    D->n_items_s = ns;
  }
  D->n_torder = 0;
  const int pad = cols_enabled() ? cols_pad(P->m, sizeof(R) == 8, P->tiles.ts) : 0;
  if (pad > 0 && D->M > 0) {
    const int nt = build_tile_order(D->d_tile_start, P->tiles.ntiles(), D->d_twork, D->d_chunks, D->d_tpartial,
                                    D->d_torder, D->d_tile_slot, &D->n_heavy, &D->n_single,
                                    P->stream);
    if (nt < 0) {
      set_error("building the tile order failed");
      return kHipError;
    }
    const size_t need = (size_t)nt * (size_t)pad * pad * pad;
    if (D->slots_cap < need) {
      dfree(D->d_slots);
      D->slots_cap = 0;
      D->d_slots = dalloc<Cplx<R>>(need);
      if (D->d_slots) D->slots_cap = need;
      (void)hipGetLastError();
    }
    // without slot memory the scatter takes the LDS-atomic path
    D->n_torder = D->d_slots ? nt : 0;
  }
  // the boxes serve only the FP32 matrix-core scatter and its fold on 8^3 tiles
  // whose padded subgrids are a multiple of four cells wide (m = 2, 4)
  const int pb = 8 + 2 * P->m;
  const bool want_boxes = sizeof(R) == 4 && P->tiles.ts[0] == 8 && P->tiles.ts[1] == 8 && P->tiles.ts[2] == 8 &&
                          pb <= 16 && pb % 4 == 0;
  D->boxes = want_boxes && D->n_torder > 0 &&
             build_tile_boxes(D->d_keys, D->M, P->tiles, D->d_torder, D->n_torder, D->d_box_tile, D->d_box_ord,
                              P->stream) == kOk;
  D->bin_geom = g;
  D->bin_tiles = P->tiles;
  D->bins_valid = true;
  return kOk;
}

template <class R>
static void kt_add(PlanT<R>* P, int k, int a, int b) {
  P->ktime[k] += P->evt.ms(a, b);
  P->kcount[k] += 1;
}

// Accumulates the reference's stage timers (seconds) and the per-kernel times.
// fwd: trafo (FFT passes rows, cols1, cols0; gather) or adj (cols0, cols1, rows;
// scatter).  The B stage is up to three kernels separated by events 10 and 13:
// tiled, fold (column scatter only) and direct.
template <class R>
static void finish_timers(PlanT<R>* P, double* timer, bool fwd, bool had_sort, bool had_f, bool had_b) {
  if (had_sort) {
    timer[T_SORT] += 1e-3 * P->evt.ms(0, 1);
    kt_add(P, PNFFT_KT_BINNING, 0, 1);
  }
  if (had_f) {
    const double t = 1e-3 * P->evt.ms(2, 3);
    timer[T_F] += t;
    kt_add(P, fwd ? PNFFT_KT_FFT_ROWS : PNFFT_KT_FFT_COLS0, 2, 8);
    kt_add(P, PNFFT_KT_FFT_COLS1, 8, 9);
    kt_add(P, fwd ? PNFFT_KT_FFT_COLS0 : PNFFT_KT_FFT_ROWS, 9, 3);
  }
  if (had_b) {
    const double t = 1e-3 * P->evt.ms(4, 5);
    timer[T_B] += t;
    timer[T_LOOP_B] += t;
    kt_add(P, fwd ? PNFFT_KT_GATHER_TILED : PNFFT_KT_SCATTER_TILED, 4, 10);
    if (!fwd) kt_add(P, PNFFT_KT_SCATTER_FOLD, 10, 13);
    kt_add(P, fwd ? PNFFT_KT_GATHER_DIRECT : PNFFT_KT_SCATTER_DIRECT, 13, 5);
  }
}

template <class T>
static T* grow(T*& p, size_t& cap, size_t need) {
  if (cap < need) {
    dfree(p);
    cap = 0;
    p = dalloc<T>(need);
    if (p) cap = need;
    (void)hipGetLastError();
  }
  return p;
}

// B stage arguments for geometry g: bins the nodes unless the bins of
// pnfft_precompute_psi are valid for g (device-resident x), then the tile lists
// and, with valid bins, the window tables.  false on a binning failure.
template <class R>
static bool b_args(PlanT<R>* P, NodesT<R>* D, R* dx, const GridGeom& g, bool host_x, bool for_scatter,
                   bool timed, BArgs<R>& A, bool* had_sort, bool need_grad = true) {
  // plans whose windows come from tables bin and tabulate on every call
  const bool reuse = !P->tab_only && D->bins_valid && (D->precompute_flags & fl::PRE_PSI) && !host_x &&
                     same_geom(D->bin_geom, g);
  if (!reuse) {
    if (timed) P->evt.rec(0, P->stream);
    if (bin(P, D, dx, g) != kOk) return false;
    if (timed) P->evt.rec(1, P->stream);
    if (had_sort) *had_sort = true;
  }
  A = BArgs<R>{};
  A.g = g;
  A.M = D->M;
  A.x = dx;
  A.perm = D->d_perm;
  A.grid = P->d_grid;
  A.c2r = P->c2r;
  set_tiles(A, D, for_scatter);
  if (reuse) set_tables(A, P, D);
  if (P->tab_only && D->M > 0) {
    const size_t M = (size_t)D->M, cm = 2 * (size_t)P->m + 1;
    R* psi = grow(D->d_pre_psi, D->pre_psi_cap, 3 * cm * M);
    R* dpsi = need_grad ? grow(D->d_pre_dpsi, D->pre_dpsi_cap, 3 * cm * M) : nullptr;
    R* xo = grow(D->d_pre_x, D->pre_x_cap, 3 * M);
    D->pre_valid = false;  // the precomputed tables, if any, are overwritten
    if (!psi || !xo || (need_grad && !dpsi) || ext_tables<R>(A, P->ip, psi, dpsi, xo, P->stream) != kOk) {
      set_error("window tables for the plan's window failed (device memory or launch)");
      return false;
    }
    A.pre_psi = psi;
    A.pre_dpsi = dpsi;
    A.pre_x = xo;
  }
  return true;
}

// scratch of the ik-differentiated transforms and the Hessian gather
template <class R>
static Cplx<R>* node_tmp(NodesT<R>* D) {
  return grow(D->d_tmp, D->tmp_cap, (size_t)D->M);
}

template <class R>
static int hessian_stage(PlanT<R>* P, NodesT<R>* D, BArgs<R> A, Cplx<R>* dh, int acc) {
  const size_t cm = 2 * (size_t)P->m + 1, M = (size_t)D->M;
  R* tab = grow(D->d_hess_tab, D->hess_tab_cap, 9 * cm * M);
  int* st = grow(D->d_hess_st, D->hess_st_cap, 3 * M);
  if (!tab || !st) {
    set_error("device allocation for the Hessian gather failed");
    return kHipError;
  }
  A.accumulate = acc;
  A.real_only = (P->pnfft_flags & fl::REAL_F) ? 1 : 0;
  return hessian_gather<R>(A, P->ip, dh, tab, st, P->stream);
}

// Complex factor tables of this call (ctab_fill) uploaded to the device:
// slot (pass, comp) at d_ctab + (pass * kCtabComps + comp) * 2 (N0 + N1 + N2).
template <class R>
static const double* ctab_upload(PlanT<R>* P, bool adjoint, bool omit_d, int npass) {
  const size_t per = 2 * (size_t)(P->N[0] + P->N[1] + P->N[2]);
  for (int ps = 0; ps < npass; ps++)
    for (int c = 0; c < kCtabComps; c++)
      ctab_fill(P, ps == 1, adjoint, c, omit_d, P->h_ctab + (size_t)(ps * kCtabComps + c) * per);
  (void)hipMemcpyAsync(P->d_ctab, P->h_ctab, sizeof(double) * per * npass * kCtabComps, hipMemcpyHostToDevice,
                       P->stream);
  return P->d_ctab;
}

template <class R>
static const double* ctab_slot(const PlanT<R>* P, int pass, int comp) {
  return P->d_ctab + (size_t)(pass * kCtabComps + comp) * 2 * (P->N[0] + P->N[1] + P->N[2]);
}

// Interlaced and/or ik-differentiated trafo (api/api-basic.c:100-244): per pass
// (the plain nodes, then for interlaced plans the nodes shifted by half a cell,
// kernel/ndft-parallel.c:2732-2753) F with the pass's complex factors, then B;
// with DIFF_IK one F and an f-gather per requested component (f, grad_t,
// Hessian (a, b)), the components landing in grad / hess through a strided add.
// Outputs are written by the first pass and added to by the second.
template <class R>
static bool trafo_passes(PlanT<R>* P, NodesT<R>* D, unsigned cf, R* dx, Cplx<R>* dfh, Cplx<R>* df, Cplx<R>* dg,
                         Cplx<R>* dh, bool host_x, bool accumulate) {
  hipStream_t s = P->stream;
  const bool il = P->pnfft_flags & fl::INTERLACED, ik = P->pnfft_flags & fl::DIFF_IK;
  const bool want_f = cf & fl::COMPUTE_F, want_g = cf & fl::COMPUTE_GRAD_F, want_h = cf & fl::COMPUTE_HESSIAN_F;
  const bool conv = !(cf & fl::OMIT_CONV) && D && D->M > 0;
  const int npass = il ? 2 : 1;
  const int comps = P->c2r ? 1 : 2;
  const long M = D ? D->M : 0;
  ctab_upload(P, false, cf & fl::OMIT_DECONV, npass);
  auto fwd = [&](int ps, int comp) {
    if (cf & fl::OMIT_FFT) return true;
    if (fft3d_forward<R>(P->fft, dfh, nullptr, P->d_buf1, P->d_buf2, P->d_grid, s, P->c2r, nullptr, nullptr,
                         ctab_slot(P, ps, comp)) != kOk) {
      set_error("FFT launch failed");
      return false;
    }
    return true;
  };
  for (int ps = 0; ps < npass; ps++) {
    GridGeom g = P->geom;
    g.shift = ps ? 0.5 : 0.0;
    const int acc = (accumulate || ps > 0) ? 1 : 0;
    BArgs<R> A{};
    if (conv && !b_args(P, D, dx, g, host_x, false, false, A, nullptr, !ik && want_g)) return false;
    A.real_only = (P->pnfft_flags & fl::REAL_F) ? 1 : 0;
    if (!ik) {
      if (!fwd(ps, 0)) return false;
      if (!conv) continue;
      if (want_f || want_g) {
        A.f = df;
        A.grad = dg;
        A.accumulate = acc;
        const int mode = (want_f && want_g) ? B_FGRAD : (want_f ? B_F : B_GRAD);
        if (gather<R>(A, mode, s) != kOk) {
          set_error("gather launch failed");
          return false;
        }
      }
      if (want_h && hessian_stage(P, D, A, dh, acc) != kOk) return false;
      continue;
    }
    Cplx<R>* tmp = conv && (want_g || want_h) ? node_tmp(D) : nullptr;
    if (conv && (want_g || want_h) && !tmp) {
      set_error("device allocation for ik-differentiation failed");
      return false;
    }
    // one transform and f-gather per component: (comp slot, output, stride, offset)
    auto component = [&](int comp, Cplx<R>* out, int stride, int offset) {
      if (!fwd(ps, comp)) return false;
      if (!conv) return true;
      A.f = stride == 1 ? out : tmp;
      A.grad = nullptr;
      A.accumulate = stride == 1 ? acc : 0;
      if (gather<R>(A, B_F, s) != kOk) {
        set_error("gather launch failed");
        return false;
      }
      if (stride > 1 && strided_put<R>(reinterpret_cast<R*>(out), reinterpret_cast<const R*>(tmp), M, comps, stride,
                                       offset, acc, s) != kOk) {
        set_error("strided copy launch failed");
        return false;
      }
      return true;
    };
    if (want_f && !component(0, df, 1, 0)) return false;
    if (want_g)
      for (int t = 0; t < 3; t++)
        if (!component(1 + t, dg, 3, t)) return false;
    if (want_h)
      for (int q = 0; q < 6; q++)
        if (!component(4 + q, dh, 6, q)) return false;
  }
  return true;
}

// api/api-basic.c:199-244
template <class R>
static void trafo_impl(PlanT<R>* P, NodesT<R>* D, unsigned cf) {
  clear_error();
  if (!P) return;
  if (!D && !(cf & fl::OMIT_CONV)) return;
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = P->stream;
  const long M = D ? D->M : 0;
  const bool want_f = cf & fl::COMPUTE_F, want_g = cf & fl::COMPUTE_GRAD_F, want_h = cf & fl::COMPUTE_HESSIAN_F;
  const bool accumulate = (cf & fl::COMPUTE_ACCUMULATED) != 0;
  bool host_x = false, host_f = false, host_g = false, host_h = false, host_fh = false;

  // c2r plans keep f, grad_f and hessian_f as reals (kernel/ndft-parallel.c:2826-2847)
  const size_t vsz = P->c2r ? sizeof(R) : sizeof(Cplx<R>);
  R* dx = D ? stage_in<R>(D->x, D->d_x, 3 * M, true, s, &host_x) : nullptr;
  Cplx<R>* dfh = stage_in<Cplx<R>>(P->f_hat, P->d_fhat, P->N_total, true, s, &host_fh,
                                   sizeof(Cplx<R>) * P->fhat_elems);
  const bool direct = cf & fl::COMPUTE_DIRECT;
  // outputs: uploaded only when accumulating (the NDFT overwrites them regardless, :400-414)
  const bool up = accumulate && !direct;
  Cplx<R>* df = nullptr;
  Cplx<R>* dg = nullptr;
  Cplx<R>* dh = nullptr;
  if (D && want_f) df = stage_in<Cplx<R>>(D->f, D->d_f, M, up, s, &host_f, vsz * M);
  if (D && want_g) dg = stage_in<Cplx<R>>(D->grad_f, D->d_grad, 3 * M, up, s, &host_g, vsz * 3 * M);
  if (D && want_h) dh = stage_in<Cplx<R>>(D->hessian_f, D->d_hess, 6 * M, up, s, &host_h, vsz * 6 * M);
  if ((want_f && M > 0 && !df) || (want_g && M > 0 && !dg) || (want_h && M > 0 && !dh) ||
      (M > 0 && !dx && !(cf & fl::OMIT_CONV)) || !dfh) {
    set_error("trafo: missing node or f_hat arrays");
    return;
  }

  bool had_sort = false, had_f = false, had_b = false;
  const bool general = P->pnfft_flags & (fl::INTERLACED | fl::DIFF_IK);
  if (direct) {
    NdftArgs<R> A{};
    for (int t = 0; t < 3; t++) A.N[t] = (int)P->N[t];
    A.M = M;
    A.x = dx;
    A.fhat = dfh;
    A.f = df;
    A.grad = dg;
    A.hess = dh;
    A.want_f = want_f;
    A.want_grad = want_g;
    A.want_hess = want_h;
    A.c2r = P->c2r;
    if (ndft_trafo<R>(A, s) != kOk) set_error("NDFT trafo launch failed");
  } else if (general) {
    if (!trafo_passes(P, D, cf, dx, dfh, df, dg, dh, host_x, accumulate)) return;
  } else {
    // D and F (fused): grid <- F D f_hat
    if (!(cf & fl::OMIT_FFT)) {
      P->evt.rec(2, s);
      const double* tab = (cf & fl::OMIT_DECONV) ? P->d_tab_scale : P->d_tab;
      if (fft3d_forward<R>(P->fft, dfh, tab, P->d_buf1, P->d_buf2, P->d_grid, s, P->c2r, P->evt.mark(8),
                           P->evt.mark(9)) != kOk)
        set_error("FFT launch failed");
      P->evt.rec(3, s);
      had_f = true;
    }
    // B
    if (!(cf & fl::OMIT_CONV) && M > 0 && (want_f || want_g || want_h)) {
      BArgs<R> A{};
      if (!b_args(P, D, dx, P->geom, host_x, false, true, A, &had_sort, want_g)) return;
      A.f = df;
      A.grad = dg;
      A.accumulate = accumulate;
      A.real_only = (P->pnfft_flags & fl::REAL_F) ? 1 : 0;
      A.mid = P->evt.mark(10);
      A.mid2 = P->evt.mark(13);
      P->evt.rec(4, s);
      if (want_f || want_g) {
        const int mode = (want_f && want_g) ? B_FGRAD : (want_f ? B_F : B_GRAD);
        if (gather<R>(A, mode, s) != kOk) set_error("gather launch failed");
      } else {
        P->evt.rec(10, s);
        P->evt.rec(13, s);
      }
      P->evt.rec(5, s);
      had_b = true;
      if (want_h && hessian_stage(P, D, A, dh, accumulate) != kOk) return;
    }
  }
  if (host_f) (void)hipMemcpyAsync(D->f, df, vsz * M, hipMemcpyDeviceToHost, s);
  if (host_g) (void)hipMemcpyAsync(D->grad_f, dg, vsz * 3 * M, hipMemcpyDeviceToHost, s);
  if (host_h) (void)hipMemcpyAsync(D->hessian_f, dh, vsz * 6 * M, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("trafo: HIP error %s", hipGetErrorString(hipGetLastError()));
    return;
  }
  finish_timers(P, P->timer_trafo, true, had_sort, had_f, had_b);
  P->timer_trafo[T_ITER] += 1;
  P->timer_trafo[T_WHOLE] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Interlaced and/or ik-differentiated adjoint (api/api-basic.c:247-378): per
// pass, B^T of the node values and F^H with the pass's complex factors (D^T,
// the conjugate interlacing phase, +2 pi i k_t for gradient components), each
// transform adding onto f_hat.
template <class R>
static bool adj_passes(PlanT<R>* P, NodesT<R>* D, unsigned cf, R* dx, Cplx<R>* dfh, Cplx<R>* df, Cplx<R>* dg,
                       bool host_x) {
  hipStream_t s = P->stream;
  const bool il = P->pnfft_flags & fl::INTERLACED, ik = P->pnfft_flags & fl::DIFF_IK;
  const bool use_f = cf & fl::COMPUTE_F, use_g = cf & fl::COMPUTE_GRAD_F;
  const bool conv = !(cf & fl::OMIT_CONV) && D && D->M > 0;
  const bool back = !(cf & fl::OMIT_FFT) && !(cf & fl::OMIT_DECONV);
  const int npass = il ? 2 : 1;
  const int comps = P->c2r ? 1 : 2;
  const long M = D ? D->M : 0;
  ctab_upload(P, true, false, npass);
  for (int ps = 0; ps < npass; ps++) {
    GridGeom g = P->geom;
    g.shift = ps ? 0.5 : 0.0;
    BArgs<R> A{};
    if (conv && (use_f || use_g) && !b_args(P, D, dx, g, host_x, true, false, A, nullptr, !ik && use_g)) return false;
    // grid <- B^T (values) for one mode; then f_hat += factors(comp) F^H grid
    auto stage = [&](int mode, Cplx<R>* f, Cplx<R>* gr, int comp) {
      if (!(cf & fl::OMIT_CONV)) {
        if (conv) {
          A.f = f;
          A.grad = gr;
          if (scatter<R>(A, mode, s) != kOk) {
            set_error("scatter launch failed");
            return false;
          }
        } else {
          fill_zero<R>(P->d_grid, (long)P->fft.grid_elems(), s);
        }
      }
      if (back && fft3d_backward<R>(P->fft, P->d_grid, nullptr, P->d_buf1, P->d_buf2, dfh, 1, s, P->c2r, nullptr,
                                    nullptr, ctab_slot(P, ps, comp)) != kOk) {
        set_error("FFT launch failed");
        return false;
      }
      return true;
    };
    if (!ik) {
      if (use_f || use_g) {
        if (!stage((use_f && use_g) ? B_FGRAD : (use_f ? B_F : B_GRAD), df, dg, 0)) return false;
      }
      continue;
    }
    if (use_f && !stage(B_F, df, nullptr, 0)) return false;
    if (use_g) {
      Cplx<R>* tmp = conv ? node_tmp(D) : nullptr;
      if (conv && !tmp) {
        set_error("device allocation for ik-differentiation failed");
        return false;
      }
      for (int t = 0; t < 3; t++) {
        if (conv && strided_get<R>(reinterpret_cast<R*>(tmp), reinterpret_cast<const R*>(dg), M, comps, 3, t, s) != kOk) {
          set_error("strided copy launch failed");
          return false;
        }
        if (!stage(B_F, tmp, nullptr, 1 + t)) return false;
      }
    }
  }
  return true;
}

// api/api-basic.c:344-378
template <class R>
static void adj_impl(PlanT<R>* P, NodesT<R>* D, unsigned cf) {
  clear_error();
  if (!P) return;
  if (!D && !(cf & fl::OMIT_CONV)) return;
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = P->stream;
  const long M = D ? D->M : 0;
  const bool use_f = cf & fl::COMPUTE_F, use_g = cf & fl::COMPUTE_GRAD_F;
  const bool accumulate = (cf & fl::COMPUTE_ACCUMULATED) != 0;
  bool host_x = false, host_f = false, host_g = false, host_fh = false;
  const size_t vsz = P->c2r ? sizeof(R) : sizeof(Cplx<R>);
  const size_t fh_bytes = sizeof(Cplx<R>) * P->fhat_elems;
  R* dx = D ? stage_in<R>(D->x, D->d_x, 3 * M, true, s, &host_x) : nullptr;
  Cplx<R>* df = (D && use_f) ? stage_in<Cplx<R>>(D->f, D->d_f, M, true, s, &host_f, vsz * M) : nullptr;
  Cplx<R>* dg =
      (D && use_g) ? stage_in<Cplx<R>>(D->grad_f, D->d_grad, 3 * M, true, s, &host_g, vsz * 3 * M) : nullptr;
  Cplx<R>* dfh = stage_in<Cplx<R>>(P->f_hat, P->d_fhat, P->N_total, accumulate, s, &host_fh, fh_bytes);
  if (!dfh || (M > 0 && !(cf & fl::OMIT_CONV) && (!dx || (use_f && !df) || (use_g && !dg)))) {
    set_error("adj: missing node or f_hat arrays");
    return;
  }
  // zero f_hat unless accumulating (:355-356)
  if (!accumulate) fill_zero<R>(dfh, P->fhat_elems, s);

  bool had_sort = false, had_f = false, had_b = false;
  if (cf & fl::COMPUTE_DIRECT) {
    NdftArgs<R> A{};
    for (int t = 0; t < 3; t++) A.N[t] = (int)P->N[t];
    A.M = M;
    A.x = dx;
    A.fhat = dfh;
    A.f = df;
    A.grad = dg;
    A.want_f = use_f;
    A.want_grad = use_g;
    A.c2r = P->c2r;
    if (ndft_adj<R>(A, s) != kOk) set_error("NDFT adj launch failed");
  } else if (P->pnfft_flags & (fl::INTERLACED | fl::DIFF_IK)) {
    if (!adj_passes(P, D, cf, dx, dfh, df, dg, host_x)) return;
  } else {
    if (!(cf & fl::OMIT_CONV)) {
      // grid = B^T (f, grad): the reference scatters into the zeroed grid (:2629-2635);
      // scatter() writes every cell itself
      if (M > 0 && (use_f || use_g)) {
        BArgs<R> A{};
        if (!b_args(P, D, dx, P->geom, host_x, true, true, A, &had_sort, use_g)) return;
        A.f = df;
        A.grad = dg;
        A.mid = P->evt.mark(10);
        A.mid2 = P->evt.mark(13);
        const int mode = (use_f && use_g) ? B_FGRAD : (use_f ? B_F : B_GRAD);
        P->evt.rec(4, s);
        if (scatter<R>(A, mode, s) != kOk) set_error("scatter launch failed");
        P->evt.rec(5, s);
        had_b = true;
      } else {
        fill_zero<R>(P->d_grid, (long)P->fft.grid_elems(), s);
      }
    }
    // F^H and D^T (fused), f_hat += D^T F^H grid
    if (!(cf & fl::OMIT_FFT) && !(cf & fl::OMIT_DECONV)) {
      P->evt.rec(2, s);
      if (fft3d_backward<R>(P->fft, P->d_grid, P->d_tab, P->d_buf1, P->d_buf2, dfh, 1, s, P->c2r,
                            P->evt.mark(8), P->evt.mark(9)) != kOk)
        set_error("FFT launch failed");
      P->evt.rec(3, s);
      had_f = true;
    }
  }
  if (host_fh) (void)hipMemcpyAsync(P->f_hat, dfh, fh_bytes, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("adj: HIP error %s", hipGetErrorString(hipGetLastError()));
    return;
  }
  finish_timers(P, P->timer_adj, false, had_sort, had_f, had_b);
  P->timer_adj[T_ITER] += 1;
  P->timer_adj[T_WHOLE] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// x-adjoint of f = trafo(f_hat): grad_x[3j+t] = Re(conj(w_j) df_j/dx_t)
template <class R>
static void vjp_impl(PlanT<R>* P, NodesT<R>* D, const Cplx<R>* w, R* gx) {
  clear_error();
  if (!P || !D) return;
  if ((P->pnfft_flags & fl::INTERLACED) || P->tab_only) {
    set_error("trafo_grad_x_vjp: interlaced plans and plans with the Bessel-I0 or interpolated windows are not "
              "supported");
    return;
  }
  hipStream_t s = P->stream;
  const long M = D->M;
  if (M == 0) return;
  bool host_x = false, host_fh = false, host_w = false, host_gx = false;
  R* dx = stage_in<R>(D->x, D->d_x, 3 * M, true, s, &host_x);
  Cplx<R>* dfh = stage_in<Cplx<R>>(P->f_hat, P->d_fhat, P->N_total, true, s, &host_fh,
                                   sizeof(Cplx<R>) * P->fhat_elems);
  const size_t vsz = P->c2r ? sizeof(R) : sizeof(Cplx<R>);
  Cplx<R>* dw = stage_in<Cplx<R>>((void*)w, D->d_f, M, true, s, &host_w, vsz * M);
  R* dgx = stage_in<R>(gx, D->d_gx, 3 * M, false, s, &host_gx);
  if (!dx || !dfh || !dw || !dgx) {
    set_error("trafo_grad_x_vjp: missing arrays");
    return;
  }
  if (fft3d_forward<R>(P->fft, dfh, P->d_tab, P->d_buf1, P->d_buf2, P->d_grid, s, P->c2r) != kOk) {
    set_error("FFT launch failed");
    return;
  }
  const bool reuse = D->bins_valid && (D->precompute_flags & fl::PRE_PSI) && !host_x &&
                     same_geom(D->bin_geom, P->geom);
  if (!reuse && bin(P, D, dx, P->geom) != kOk) return;
  BArgs<R> A{};
  A.g = P->geom;
  A.M = M;
  A.x = dx;
  A.perm = D->d_perm;
  A.grid = P->d_grid;
  A.w = dw;
  A.gx = dgx;
  A.c2r = P->c2r;
  set_tiles(A, D);
  if (reuse) set_tables(A, P, D);
  if (gather<R>(A, B_VJP, s) != kOk) set_error("gather launch failed");
  if (host_gx) (void)hipMemcpyAsync(gx, dgx, sizeof(R) * 3 * M, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) set_error("vjp: HIP error %s", hipGetErrorString(hipGetLastError()));
}

static bool same_full_geom(const GridGeom& a, const GridGeom& b) {
  for (int t = 0; t < 3; t++)
    if (a.n[t] != b.n[t] || a.b[t] != b.b[t] || a.ws[t] != b.ws[t]) return false;
  return a.m == b.m && a.win == b.win && a.shift == b.shift;
}


// kernel/ndft-parallel.c:1144-1245: bins the nodes and, for PNFFT_PRE_PSI,
// tabulates their window factors (PNFFT_PRE_GRAD_PSI: also the derivatives) in
// the bin order; set_x discards both.  Without device memory for the tables the
// B kernels evaluate the windows as without PRE_PSI.
template <class R>
static void precompute_impl(PlanT<R>* P, NodesT<R>* D, unsigned flags) {
  clear_error();
  if (!P || !D) return;
  D->precompute_flags = flags;
  D->pre_valid = false;
  if (!(flags & fl::PRE_PSI) || D->M == 0) return;
  bool host_x = false;
  R* dx = stage_in<R>(D->x, D->d_x, 3 * D->M, true, P->stream, &host_x);
  if (!dx) return;
  if (bin(P, D, dx, P->geom) != kOk) return;
  // plans whose windows come from tables tabulate on every call (b_args)
  if (!host_x && P->m >= 2 && P->m <= 8 && tables_enabled() && !P->tab_only) {
    const size_t M = (size_t)D->M, cm = 2 * (size_t)P->m + 1;
    const bool want_grad = (flags & fl::PRE_GRAD_PSI) != 0;
    R* psi = grow(D->d_pre_psi, D->pre_psi_cap, 3 * cm * M);
    R* dpsi = want_grad ? grow(D->d_pre_dpsi, D->pre_dpsi_cap, 3 * cm * M) : nullptr;
    R* xo = grow(D->d_pre_x, D->pre_x_cap, 3 * M);
    if (psi && xo && (dpsi || !want_grad)) {
      BArgs<R> A{};
      A.g = P->geom;
      A.M = D->M;
      A.x = dx;
      A.perm = D->d_perm;
      if (precompute_tables<R>(A, psi, dpsi, xo, P->stream) == kOk) {
        D->pre_valid = true;
        D->pre_grad = dpsi != nullptr;
        D->pre_geom = P->geom;
      }
    }
  }
  (void)hipStreamSynchronize(P->stream);
  // host x is re-uploaded on every call, so its bins are not reused
  if (host_x) D->bins_valid = false;
}

// window tables for a B stage that reuses the precomputed bins
template <class R>
static void set_tables(BArgs<R>& A, const PlanT<R>* P, const NodesT<R>* D) {
  (void)P;
  if (!D->pre_valid || !same_full_geom(D->pre_geom, A.g)) return;
  A.pre_psi = D->d_pre_psi;
  A.pre_dpsi = D->pre_grad ? D->d_pre_dpsi : nullptr;
  A.pre_x = D->d_pre_x;
}

template <class R>
static int node_bins_impl(PlanT<R>* P, NodesT<R>* D, int* order, unsigned* keys) {
  clear_error();
  if (!P || !D) return -1;
  if (D->M == 0) return 0;
  bool host_x = false;
  R* dx = stage_in<R>(D->x, D->d_x, 3 * D->M, true, P->stream, &host_x);
  if (!dx) return -1;
  if (!(D->bins_valid && !host_x && same_geom(D->bin_geom, P->geom)))
    if (bin(P, D, dx, P->geom) != kOk) return -1;
  if (order) (void)hipMemcpyAsync(order, D->d_perm, sizeof(int) * D->M, hipMemcpyDefault, P->stream);
  if (keys) (void)hipMemcpyAsync(keys, D->d_keys, sizeof(unsigned) * D->M, hipMemcpyDefault, P->stream);
  if (hipStreamSynchronize(P->stream) != hipSuccess) return -1;
  if (host_x) D->bins_valid = false;
  return 0;
}

// node_bins without allocation: x, order and keys are device arrays of the
// caller; the sort runs in the library's static scratch (bin_nodes_noalloc).
template <class R>
static int node_bins_noalloc_impl(PlanT<R>* P, const R* x, long M, int* order, unsigned* keys) {
  clear_error();
  if (!P) return -1;
  if (M < 0 || M > kNoAllocMaxNodes) {
    set_error("node_bins_noalloc: M=%ld outside [0, %ld] (the static scratch holds %ld nodes)", M, kNoAllocMaxNodes,
              kNoAllocMaxNodes);
    return -1;
  }
  if (M == 0) return 0;
  if (!x || !order || !keys || !is_device_ptr(x) || !is_device_ptr(order) || !is_device_ptr(keys)) {
    set_error("node_bins_noalloc: x, order and keys must be device arrays");
    return -1;
  }
  const int st = bin_nodes_noalloc<R>(x, M, P->geom, P->tiles, order, keys, P->stream);
  if (st != kOk) {
    set_error("node_bins_noalloc: binning failed (%d)", st);
    return -1;
  }
  if (hipStreamSynchronize(P->stream) != hipSuccess) {
    set_error("node_bins_noalloc: HIP error %s", hipGetErrorString(hipGetLastError()));
    return -1;
  }
  return 0;
}

// glibc-rand based test data (api/api-basic.c:681-704, api/api-adv.c:35-86)
template <class R>
static void init_x_3d_adv(const R* lo, const R* up, const R* x_max, long M, R* x) {
  for (long j = 0; j < M; j++)
    for (int t = 0; t < 3; t++) {
      R v;
      do {
        v = (R)((double)rand() / RAND_MAX) * (up[t] - lo[t]) + lo[t];
      } while ((v < lo[t]) || (v >= up[t]) || (v < -x_max[t]) || (v >= x_max[t]));
      x[3 * j + t] = v;
    }
}

}  // namespace pnfft_amd

// ---------------------------------------------------------------------------
// C interface (include/pnfft.h), both precisions.  The opaque handles of the
// header are the plan and node objects; the header is included so that every
// definition below is checked against its declaration.
// ---------------------------------------------------------------------------
struct pnfft_plan_s : pnfft_amd::PlanT<double> {};
struct pnfftf_plan_s : pnfft_amd::PlanT<float> {};
struct pnfft_nodes_s : pnfft_amd::NodesT<double> {};
struct pnfftf_nodes_s : pnfft_amd::NodesT<float> {};

namespace pnfft_amd {
template <>
struct Handles<double> {
  using plan = pnfft_plan_s;
  using nodes = pnfft_nodes_s;
};
template <>
struct Handles<float> {
  using plan = pnfftf_plan_s;
  using nodes = pnfftf_nodes_s;
};
}  // namespace pnfft_amd

using namespace pnfft_amd;

#define PNFFT_DEFINE_IMPL(PREFIX, R, CT)                                                                   \
  static inline PlanT<R>* PREFIX##P(PREFIX##plan p) { return p; }                                          \
  static inline NodesT<R>* PREFIX##D(PREFIX##nodes h) { return h; }                                        \
  extern "C" {                                                                                             \
  int PREFIX##create_procmesh_2d(MPI_Comm comm, int np0, int np1, MPI_Comm* cc) {                           \
    (void)np0; (void)np1; if (cc) *cc = comm; return 0;                                                     \
  }                                                                                                        \
  int PREFIX##create_procmesh(int rnk, MPI_Comm comm, const int* np, MPI_Comm* cc) {                        \
    (void)rnk; (void)np; if (cc) *cc = comm; return 0;                                                      \
  }                                                                                                        \
  static void PREFIX##local_sizes(const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m,           \
                                  ptrdiff_t* lN, ptrdiff_t* lNs, R* lo, R* up) {                            \
    for (int t = 0; t < 3; t++) {                                                                          \
      lN[t] = N[t]; lNs[t] = -N[t] / 2;                                                                     \
      const long c = (long)std::lrint(std::floor((double)n[t] * (double)x_max[t]));                         \
      const long no = std::min<long>(n[t], 2 * (c + m + 2));                                                \
      /* node_borders, kernel/ndft-parallel.c:734-775 */                                                    \
      double l = (double)(-no / 2) / (double)n[t], u = (double)(-no / 2 + no) / (double)n[t];              \
      if (l < -x_max[t]) l = -x_max[t]; if (l > x_max[t]) l = x_max[t];                                     \
      if (u < -x_max[t]) u = -x_max[t]; if (u > x_max[t]) u = x_max[t];                                     \
      lo[t] = (R)l; up[t] = (R)u;                                                                           \
    }                                                                                                      \
  }                                                                                                        \
  void PREFIX##local_size_guru(int d, const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m,       \
                               MPI_Comm cc, unsigned pf, ptrdiff_t* lN, ptrdiff_t* lNs, R* lo, R* up) {     \
    (void)cc; (void)pf;                                                                                     \
    if (d != 3) { set_error("d != 3 not yet implemented"); return; }                                        \
    PREFIX##local_sizes(N, n, x_max, m, lN, lNs, lo, up);                                                   \
  }                                                                                                        \
  void PREFIX##local_size_guru_c2r(int d, const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m,   \
                                   MPI_Comm cc, unsigned pf, ptrdiff_t* lN, ptrdiff_t* lNs, R* lo, R* up) { \
    PREFIX##local_size_guru(d, N, n, x_max, m, cc, pf, lN, lNs, lo, up);                                    \
    if (d == 3) lN[2] = N[2] / 2 + 1;                                                                       \
  }                                                                                                        \
  void PREFIX##local_size_adv(int d, const ptrdiff_t* N, MPI_Comm cc, unsigned pf, ptrdiff_t* lN,          \
                              ptrdiff_t* lNs, R* lo, R* up) {                                               \
    if (d != 3) { set_error("d != 3 not yet implemented"); return; }                                        \
    ptrdiff_t n[3] = {2 * N[0], 2 * N[1], 2 * N[2]}; R xm[3] = {R(0.5), R(0.5), R(0.5)};                    \
    PREFIX##local_size_guru(d, N, n, xm, 6, cc, pf, lN, lNs, lo, up);                                       \
  }                                                                                                        \
  void PREFIX##local_size_adv_c2r(int d, const ptrdiff_t* N, MPI_Comm cc, unsigned pf, ptrdiff_t* lN,      \
                                  ptrdiff_t* lNs, R* lo, R* up) {                                           \
    if (d != 3) { set_error("d != 3 not yet implemented"); return; }                                        \
    ptrdiff_t n[3] = {2 * N[0], 2 * N[1], 2 * N[2]}; R xm[3] = {R(0.5), R(0.5), R(0.5)};                    \
    PREFIX##local_size_guru_c2r(d, N, n, xm, 6, cc, pf, lN, lNs, lo, up);                                   \
  }                                                                                                        \
  void PREFIX##local_size_3d(const ptrdiff_t* N, MPI_Comm cc, unsigned pf, ptrdiff_t* lN, ptrdiff_t* lNs,  \
                             R* lo, R* up) { PREFIX##local_size_adv(3, N, cc, pf, lN, lNs, lo, up); }        \
  void PREFIX##local_size_3d_c2r(const ptrdiff_t* N, MPI_Comm cc, unsigned pf, ptrdiff_t* lN,              \
                                 ptrdiff_t* lNs, R* lo, R* up) {                                            \
    PREFIX##local_size_adv_c2r(3, N, cc, pf, lN, lNs, lo, up);                                              \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_guru(int d, const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m,     \
                                 unsigned pnfft_flags, unsigned pfft_flags, MPI_Comm cc) {                  \
    (void)cc;                                                                                               \
    return static_cast<PREFIX##plan>(plan_create<R>(d, N, n, x_max, m, pnfft_flags, pfft_flags, false));    \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_guru_c2r(int d, const ptrdiff_t* N, const ptrdiff_t* n, const R* x_max, int m, \
                                     unsigned pnfft_flags, unsigned pfft_flags, MPI_Comm cc) {              \
    (void)cc;                                                                                               \
    return static_cast<PREFIX##plan>(plan_create<R>(d, N, n, x_max, m, pnfft_flags, pfft_flags, true));     \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_adv(int d, const ptrdiff_t* N, unsigned pnfft_flags, unsigned pfft_flags,      \
                                MPI_Comm cc) {                                                              \
    if (d != 3) { set_error("d != 3 not yet implemented"); return nullptr; }                                \
    ptrdiff_t n[3] = {2 * N[0], 2 * N[1], 2 * N[2]}; R xm[3] = {R(0.5), R(0.5), R(0.5)};                    \
    return PREFIX##init_guru(d, N, n, xm, 6, pnfft_flags, pfft_flags, cc);                                  \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_adv_c2r(int d, const ptrdiff_t* N, unsigned pnfft_flags, unsigned pfft_flags,  \
                                    MPI_Comm cc) {                                                          \
    if (d != 3) { set_error("d != 3 not yet implemented"); return nullptr; }                                \
    ptrdiff_t n[3] = {2 * N[0], 2 * N[1], 2 * N[2]}; R xm[3] = {R(0.5), R(0.5), R(0.5)};                    \
    return PREFIX##init_guru_c2r(d, N, n, xm, 6, pnfft_flags, pfft_flags, cc);                              \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_3d(const ptrdiff_t* N, MPI_Comm cc) {                                          \
    return PREFIX##init_adv(3, N, fl::MALLOC_F_HAT, 0, cc);                                                 \
  }                                                                                                        \
  PREFIX##plan PREFIX##init_3d_c2r(const ptrdiff_t* N, MPI_Comm cc) {                                      \
    return PREFIX##init_adv_c2r(3, N, fl::MALLOC_F_HAT, 0, cc);                                             \
  }                                                                                                        \
  PREFIX##nodes PREFIX##init_nodes(ptrdiff_t M, unsigned mf) {                                              \
    return static_cast<PREFIX##nodes>(nodes_create<R>(M, mf, false));                                       \
  }                                                                                                        \
  PREFIX##nodes PREFIX##init_nodes_device(ptrdiff_t M, unsigned mf) {                                       \
    return static_cast<PREFIX##nodes>(nodes_create<R>(M, mf, true));                                        \
  }                                                                                                        \
  void PREFIX##free_nodes(PREFIX##nodes h, unsigned ff) {                                                   \
    if (!h) return;                                                                                        \
    (void)hipDeviceSynchronize();                                                                           \
    if (ff & fl::MALLOC_GRAD_F) mem_free(h->grad_f);                                                        \
    if (ff & fl::MALLOC_HESSIAN_F) mem_free(h->hessian_f);                                                  \
    if (ff & fl::MALLOC_F) mem_free(h->f);                                                                  \
    if (ff & fl::MALLOC_X) mem_free(h->x);                                                                  \
    nodes_destroy<R>(h);                                                                                    \
  }                                                                                                        \
  void PREFIX##precompute_psi(PREFIX##plan p, PREFIX##nodes h, unsigned pf) {                               \
    precompute_impl<R>(PREFIX##P(p), PREFIX##D(h), pf);                                                     \
  }                                                                                                        \
  void PREFIX##precompute_one_psi(PREFIX##plan p, PREFIX##nodes h, unsigned pf) {                           \
    PREFIX##precompute_psi(p, h, pf);                                                                       \
  }                                                                                                        \
  void PREFIX##set_f(CT* f, PREFIX##nodes h) { if (h) h->f = (R*)f; }                                       \
  void PREFIX##set_grad_f(CT* g, PREFIX##nodes h) { if (h) h->grad_f = (R*)g; }                             \
  void PREFIX##set_hessian_f(CT* g, PREFIX##nodes h) { if (h) h->hessian_f = (R*)g; }                       \
  void PREFIX##set_f_real(R* f, PREFIX##nodes h) { if (h) h->f = f; }                                       \
  void PREFIX##set_grad_f_real(R* g, PREFIX##nodes h) { if (h) h->grad_f = g; }                             \
  void PREFIX##set_hessian_f_real(R* g, PREFIX##nodes h) { if (h) h->hessian_f = g; }                       \
  void PREFIX##set_x(R* x, PREFIX##nodes h) {                                                               \
    if (!h) return;                                                                                        \
    h->x = x; h->bins_valid = false; h->pre_valid = false;                                                  \
  }                                                                                                        \
  void PREFIX##set_f_hat(CT* fh, PREFIX##plan p) { if (p) p->f_hat = fh; }                                  \
  void PREFIX##set_f_hat_real(R* fh, PREFIX##plan p) { if (p) p->f_hat = fh; }                              \
  void PREFIX##set_b(R b0, R b1, R b2, PREFIX##plan p) {                                                    \
    if (!p) return;                                                                                        \
    p->b[0] = b0; p->b[1] = b1; p->b[2] = b2;                                                               \
    for (int t = 0; t < 3; t++) p->geom.b[t] = p->b[t];                                                     \
    tables_compute(PREFIX##P(p));                                                                           \
  }                                                                                                        \
  CT* PREFIX##get_f(const PREFIX##nodes h) { return h ? (CT*)h->f : nullptr; }                           \
  CT* PREFIX##get_grad_f(const PREFIX##nodes h) { return h ? (CT*)h->grad_f : nullptr; }                 \
  CT* PREFIX##get_hessian_f(const PREFIX##nodes h) { return h ? (CT*)h->hessian_f : nullptr; }           \
  R* PREFIX##get_f_real(const PREFIX##nodes h) { return h ? h->f : nullptr; }                            \
  R* PREFIX##get_grad_f_real(const PREFIX##nodes h) { return h ? h->grad_f : nullptr; }                  \
  R* PREFIX##get_hessian_f_real(const PREFIX##nodes h) { return h ? h->hessian_f : nullptr; }            \
  R* PREFIX##get_x(const PREFIX##nodes h) { return h ? h->x : nullptr; }                                 \
  CT* PREFIX##get_f_hat(const PREFIX##plan p) { return p ? (CT*)p->f_hat : nullptr; }                    \
  R* PREFIX##get_f_hat_real(const PREFIX##plan p) { return p ? (R*)p->f_hat : nullptr; }                 \
  int PREFIX##get_d(const PREFIX##plan p) { return p->d; }                                               \
  int PREFIX##get_m(const PREFIX##plan p) { return p->m; }                                               \
  void PREFIX##get_x_max(const PREFIX##plan p, R* xm) {                                                  \
    for (int t = 0; t < 3; t++) xm[t] = (R)p->x_max[t];                                                     \
  }                                                                                                        \
  void PREFIX##get_N(const PREFIX##plan p, ptrdiff_t* N) {                                               \
    for (int t = 0; t < 3; t++) N[t] = p->N[t];                                                             \
  }                                                                                                        \
  void PREFIX##get_n(const PREFIX##plan p, ptrdiff_t* n) {                                               \
    for (int t = 0; t < 3; t++) n[t] = p->n[t];                                                             \
  }                                                                                                        \
  unsigned PREFIX##get_pnfft_flags(const PREFIX##plan p) { return p->pnfft_flags; }                      \
  unsigned PREFIX##get_pfft_flags(const PREFIX##plan p) { return p->pfft_flags; }                        \
  void PREFIX##get_b(const PREFIX##plan p, R* b0, R* b1, R* b2) {                                        \
    *b0 = (R)p->b[0]; *b1 = (R)p->b[1]; *b2 = (R)p->b[2];                                                   \
  }                                                                                                        \
  void PREFIX##finalize(PREFIX##plan p, unsigned ff) {                                                      \
    if (!p) return;                                                                                        \
    if ((ff & fl::MALLOC_F_HAT) && p->f_hat) mem_free(p->f_hat);                                            \
    plan_destroy<R>(p);                                                                                     \
  }                                                                                                        \
  void PREFIX##trafo(PREFIX##plan p, PREFIX##nodes h, unsigned cf) { trafo_impl<R>(p, h, cf); }             \
  void PREFIX##adj(PREFIX##plan p, PREFIX##nodes h, unsigned cf) { adj_impl<R>(p, h, cf); }                 \
  void PREFIX##trafo_grad_x_vjp(PREFIX##plan p, PREFIX##nodes h, const CT* w, R* gx) {                      \
    vjp_impl<R>(p, h, (const Cplx<R>*)w, gx);                                                               \
  }                                                                                                        \
  int PREFIX##node_bins(PREFIX##plan p, PREFIX##nodes h, int* order, unsigned* keys) {                      \
    return node_bins_impl<R>(p, h, order, keys);                                                            \
  }                                                                                                        \
  int PREFIX##node_bins_noalloc(PREFIX##plan p, const R* x, ptrdiff_t M, int* order, unsigned* keys) {       \
    return node_bins_noalloc_impl<R>(p, x, (long)M, order, keys);                                           \
  }                                                                                                        \
  CT* PREFIX##oversampled_grid(PREFIX##plan p) { return p ? (CT*)p->d_grid : nullptr; }                     \
  void PREFIX##init(void) {}                                                                                \
  void PREFIX##cleanup(void) {}                                                                             \
  void* PREFIX##malloc(size_t n) { return host_alloc(n); }                                                  \
  R* PREFIX##alloc_real(size_t n) { return (R*)host_alloc(sizeof(R) * n); }                                  \
  CT* PREFIX##alloc_complex(size_t n) { return (CT*)host_alloc(2 * sizeof(R) * n); }                         \
  void PREFIX##free(void* ptr) { mem_free(ptr); }                                                           \
  void* PREFIX##malloc_device(size_t n) { return device_alloc(n); }                                         \
  void PREFIX##init_f_hat_3d(const ptrdiff_t* N, const ptrdiff_t* lN, const ptrdiff_t* lNs, unsigned pf,    \
                             CT* data) {                                                                    \
    /* deterministic fill; the reference delegates to PFFT's init_input_complex_3d */                      \
    (void)pf; R* d = (R*)data; long k = 0;                                                                  \
    for (long k0 = lNs[0]; k0 < lNs[0] + lN[0]; k0++)                                                       \
      for (long k1 = lNs[1]; k1 < lNs[1] + lN[1]; k1++)                                                     \
        for (long k2 = lNs[2]; k2 < lNs[2] + lN[2]; k2++, k++) {                                            \
          const long g = ((k0 + N[0] / 2) * N[1] + (k1 + N[1] / 2)) * N[2] + (k2 + N[2] / 2);               \
          d[2 * k] = (R)(1000.0 / (2.0 * g + 1.0)); d[2 * k + 1] = (R)(1000.0 / (2.0 * g + 2.0));           \
        }                                                                                                  \
  }                                                                                                        \
  void PREFIX##init_f(ptrdiff_t M, CT* data) {                                                              \
    R* d = (R*)data;                                                                                        \
    for (long j = 0; j < M; j++) {                                                                          \
      R re = (R)(100.0 * (R)rand() / RAND_MAX); R im = (R)(100.0 * (R)rand() / RAND_MAX);                   \
      d[2 * j] = re; d[2 * j + 1] = im;                                                                     \
    }                                                                                                      \
  }                                                                                                        \
  void PREFIX##init_x_3d_adv(const R* lo, const R* up, const R* xm, ptrdiff_t M, R* x) {                    \
    init_x_3d_adv<R>(lo, up, xm, M, x);                                                                     \
  }                                                                                                        \
  void PREFIX##init_x_3d(const R* lo, const R* up, ptrdiff_t M, R* x) {                                     \
    R xm[3] = {R(0.5), R(0.5), R(0.5)}; init_x_3d_adv<R>(lo, up, xm, M, x);                                  \
  }                                                                                                        \
  void PREFIX##zero_f_hat(PREFIX##plan p) {                                                                 \
    if (!p || !p->f_hat) return;                                                                           \
    if (is_device_ptr(p->f_hat)) {                                                                          \
      fill_zero<R>((Cplx<R>*)p->f_hat, p->fhat_elems, p->stream); (void)hipStreamSynchronize(p->stream);     \
    } else memset(p->f_hat, 0, sizeof(Cplx<R>) * p->fhat_elems);                                            \
  }                                                                                                        \
  R PREFIX##inv_phi_hat(const PREFIX##plan p, int dim, ptrdiff_t k) {                                    \
    return (R)plan_inv_phi_hat(PREFIX##P(p), dim, k);                                                       \
  }                                                                                                        \
  R PREFIX##phi_hat(const PREFIX##plan p, int dim, ptrdiff_t k) {                                        \
    return (R)plan_phi_hat(PREFIX##P(p), dim, k);                                                           \
  }                                                                                                        \
  R PREFIX##psi(const PREFIX##plan p, int dim, R x) {                                                    \
    return (R)host::psi(p->win, p->n[dim], p->b[dim], p->m, (double)x);                                     \
  }                                                                                                        \
  R PREFIX##dpsi(const PREFIX##plan p, int dim, R x) {                                                   \
    return (R)host::dpsi(p->win, p->n[dim], p->b[dim], p->m, (double)x);                                    \
  }                                                                                                        \
  R PREFIX##ddpsi(const PREFIX##plan p, int dim, R x) {                                                  \
    return (R)host::ddpsi(p->win, p->n[dim], p->b[dim], p->m, (double)x);                                   \
  }                                                                                                        \
  double* PREFIX##get_timer_trafo(PREFIX##plan p) {                                                         \
    double* t = (double*)malloc(sizeof(double) * T_LEN);                                                    \
    memcpy(t, p->timer_trafo, sizeof(double) * T_LEN); return t;                                            \
  }                                                                                                        \
  double* PREFIX##get_timer_adj(PREFIX##plan p) {                                                           \
    double* t = (double*)malloc(sizeof(double) * T_LEN);                                                    \
    memcpy(t, p->timer_adj, sizeof(double) * T_LEN); return t;                                              \
  }                                                                                                        \
  void PREFIX##timer_average(double* t) {                                                                    \
    if (t[T_ITER] > 0) for (int i = 1; i < T_LEN; i++) t[i] /= t[T_ITER];                                   \
  }                                                                                                        \
  double* PREFIX##timer_copy(const double* o) {                                                              \
    double* t = (double*)malloc(sizeof(double) * T_LEN); memcpy(t, o, sizeof(double) * T_LEN); return t;    \
  }                                                                                                        \
  double* PREFIX##timer_reduce_max(MPI_Comm c, double* t) { (void)c; return PREFIX##timer_copy(t); }         \
  double* PREFIX##timer_add(const double* a, const double* b) {                                              \
    double* t = (double*)malloc(sizeof(double) * T_LEN);                                                    \
    for (int i = 0; i < T_LEN; i++) t[i] = a[i] + b[i]; return t;                                           \
  }                                                                                                        \
  void PREFIX##timer_free(double* t) { free(t); }                                                           \
  void PREFIX##reset_timer(PREFIX##plan p) {                                                                \
    for (int i = 0; i < T_LEN; i++) p->timer_trafo[i] = p->timer_adj[i] = 0;                                 \
    for (int i = 0; i < PNFFT_KT_COUNT; i++) { p->ktime[i] = 0; p->kcount[i] = 0; }                          \
  }                                                                                                        \
  int PREFIX##kernel_times(PREFIX##plan p, double* ms, long* launches, int n) {                              \
    if (!p) return PNFFT_KT_COUNT;                                                                          \
    for (int i = 0; i < n && i < PNFFT_KT_COUNT; i++) {                                                     \
      if (ms) ms[i] = p->ktime[i];                                                                          \
      if (launches) launches[i] = p->kcount[i];                                                             \
    }                                                                                                      \
    return PNFFT_KT_COUNT;                                                                                  \
  }                                                                                                        \
  static void PREFIX##print_timer(const double* tm, const char* name) {                                      \
    double t[T_LEN]; memcpy(t, tm, sizeof(t));                                                              \
    const double it = t[T_ITER] > 0 ? t[T_ITER] : 1;                                                        \
    printf("%s: iterations %g, whole %.3e s, D %.3e s, F %.3e s, B %.3e s (loop_B %.3e s, sort %.3e s)\n",   \
           name, t[T_ITER], t[T_WHOLE] / it, t[T_D] / it, t[T_F] / it, t[T_B] / it, t[T_LOOP_B] / it,       \
           t[T_SORT] / it);                                                                                 \
  }                                                                                                        \
  void PREFIX##print_average_timer(const PREFIX##plan p, MPI_Comm c) {                                    \
    (void)c;                                                                                                \
    PREFIX##print_timer(p->timer_trafo, "PNFFT trafo"); PREFIX##print_timer(p->timer_adj, "PNFFT adj");      \
  }                                                                                                        \
  void PREFIX##print_average_timer_adv(const PREFIX##plan p, MPI_Comm c) {                                \
    PREFIX##print_average_timer(p, c);                                                                      \
  }                                                                                                        \
  const char* PREFIX##last_error(void) { return g_last_error.empty() ? nullptr : g_last_error.c_str(); }     \
  }

PNFFT_DEFINE_IMPL(pnfft_, double, pnfft_complex)
PNFFT_DEFINE_IMPL(pnfftf_, float, pnfftf_complex)
