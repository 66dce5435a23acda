/*
 * pnfft.h -- C interface of the MI355X (gfx950) single-GPU NFFT library.
 *
 * Drop-in for the reference's public header (api/pnfft.h:51-418): the same
 * function names, argument meaning, flag values and opaque plan/nodes handles
 * for the double (pnfft_) and single (pnfftf_) precisions.  Each declaration
 * cites the reference definition it replaces.  Long double (pnfftl_) has no
 * GPU counterpart and is not provided.
 *
 * Communicators: every init function takes an MPI_Comm as the reference does.
 * On one GPU the communicator is accepted and ignored (one process owns the
 * whole grid).  Define PNFFT_HAVE_MPI before including this header to use the
 * real <mpi.h>; otherwise a minimal MPICH-compatible MPI_Comm is declared.
 *
 * Memory: getters return the pointers the library holds.  Arrays allocated by
 * the library (init_nodes malloc flags, PNFFT_MALLOC_F_HAT, pnfft_malloc) are
 * pinned host memory, so user code can write x and f_hat through them as with
 * the reference; trafo/adj stage them to the GPU and back.  Pointers handed to
 * the set_* functions may also be device pointers (hipMalloc): those are used
 * in place and no transfer happens (device-resident mode).
 */
#ifndef PNFFT_H
#define PNFFT_H 1

#include <stddef.h>

#ifdef PNFFT_HAVE_MPI
#  include <mpi.h>
#else
#  ifndef MPI_COMM_WORLD
typedef int MPI_Comm;
#    define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#    define MPI_COMM_SELF  ((MPI_Comm)0x44000001)
#  endif
#endif

/* PFFT planner flags: accepted and ignored (there is no PFFT underneath). */
#ifndef PFFT_ESTIMATE
#  define PFFT_MEASURE         (0U)
#  define PFFT_ESTIMATE        (1U << 3)
#  define PFFT_PATIENT         (1U << 4)
#  define PFFT_EXHAUSTIVE      (1U << 5)
#  define PFFT_NO_TUNE         (0U)
#  define PFFT_TUNE            (1U << 6)
#  define PFFT_PRESERVE_INPUT  (1U << 7)
#  define PFFT_DESTROY_INPUT   (1U << 8)
#endif

/* Complex types: C99 _Complex when <complex.h> was included first (as in the
 * reference's own translation units, api/api-basic.c:21), else double[2]. */
#if defined(_Complex_I) && defined(complex) && defined(I) && !defined(__cplusplus)
typedef double _Complex pnfft_complex;
typedef float _Complex pnfftf_complex;
#else
typedef double pnfft_complex[2];
typedef float pnfftf_complex[2];
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PNFFT_DECLARE_API(PNX, R, C, INT)                                                     \
  typedef struct PNX(plan_s) *PNX(plan);                                                      \
  typedef struct PNX(nodes_s) *PNX(nodes);                                                    \
  /* api/pnfft.h:56-59 -- single GPU: copies comm into *comm_cart, returns 0 */               \
  int PNX(create_procmesh_2d)(MPI_Comm comm, int np0, int np1, MPI_Comm *comm_cart_2d);       \
  int PNX(create_procmesh)(int rnk, MPI_Comm comm, const int *np, MPI_Comm *comm_cart);       \
  /* api/pnfft.h:61-90; api/api-basic.c:36-63, api/api-adv.c:88-129, api/api-guru.c:42-107 */ \
  void PNX(local_size_3d)(const INT *N, MPI_Comm comm_cart, unsigned pnfft_flags,             \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  void PNX(local_size_3d_c2r)(const INT *N, MPI_Comm comm_cart, unsigned pnfft_flags,         \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  void PNX(local_size_adv)(int d, const INT *N, MPI_Comm comm_cart, unsigned pnfft_flags,     \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  void PNX(local_size_adv_c2r)(int d, const INT *N, MPI_Comm comm_cart, unsigned pnfft_flags, \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  void PNX(local_size_guru)(int d, const INT *N, const INT *n, const R *x_max, int m,         \
      MPI_Comm comm_cart, unsigned pnfft_flags,                                               \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  void PNX(local_size_guru_c2r)(int d, const INT *N, const INT *n, const R *x_max, int m,     \
      MPI_Comm comm_cart, unsigned pnfft_flags,                                               \
      INT *local_N, INT *local_N_start, R *lower_border, R *upper_border);                    \
  /* api/pnfft.h:92-113; api/api-basic.c:66-97, api/api-adv.c:132-169, api/api-guru.c:64-81 */ \
  PNX(plan) PNX(init_3d)(const INT *N, MPI_Comm comm_cart);                                   \
  PNX(plan) PNX(init_3d_c2r)(const INT *N, MPI_Comm comm_cart);                               \
  PNX(plan) PNX(init_adv)(int d, const INT *N, unsigned pnfft_flags, unsigned pfft_flags,     \
      MPI_Comm comm_cart);                                                                    \
  PNX(plan) PNX(init_adv_c2r)(int d, const INT *N, unsigned pnfft_flags, unsigned pfft_flags, \
      MPI_Comm comm_cart);                                                                    \
  PNX(plan) PNX(init_guru)(int d, const INT *N, const INT *n, const R *x_max, int m,          \
      unsigned pnfft_flags, unsigned pfft_flags, MPI_Comm comm_cart);                         \
  PNX(plan) PNX(init_guru_c2r)(int d, const INT *N, const INT *n, const R *x_max, int m,      \
      unsigned pnfft_flags, unsigned pfft_flags, MPI_Comm comm_cart);                         \
  /* api/pnfft.h:115-122; api/api-basic.c:402-447, kernel/ndft-parallel.c:1144-1250 */        \
  PNX(nodes) PNX(init_nodes)(INT local_M, unsigned malloc_flags);                             \
  void PNX(free_nodes)(PNX(nodes) nodes, unsigned pnfft_finalize_flags);                      \
  void PNX(precompute_psi)(PNX(plan) ths, PNX(nodes) nodes, unsigned precompute_flags);       \
  /* Alias of precompute_psi under the name used in the task statement. */                   \
  void PNX(precompute_one_psi)(PNX(plan) ths, PNX(nodes) nodes, unsigned precompute_flags);   \
  /* api/pnfft.h:124-182; api/api-basic.c:457-661 */                                          \
  void PNX(set_f)(C *f, PNX(nodes) nodes);                                                    \
  void PNX(set_grad_f)(C *grad_f, PNX(nodes) nodes);                                          \
  void PNX(set_hessian_f)(C *hessian_f, PNX(nodes) nodes);                                    \
  void PNX(set_f_real)(R *f, PNX(nodes) nodes);                                               \
  void PNX(set_grad_f_real)(R *grad_f, PNX(nodes) nodes);                                     \
  void PNX(set_hessian_f_real)(R *hessian_f, PNX(nodes) nodes);                               \
  void PNX(set_x)(R *x, PNX(nodes) nodes);                                                    \
  void PNX(set_f_hat)(C *f_hat, PNX(plan) ths);                                               \
  void PNX(set_f_hat_real)(R *f_hat, PNX(plan) ths);                                          \
  void PNX(set_b)(R b0, R b1, R b2, PNX(plan) ths);                                           \
  C *PNX(get_f)(const PNX(nodes) nodes);                                                      \
  C *PNX(get_grad_f)(const PNX(nodes) nodes);                                                 \
  C *PNX(get_hessian_f)(const PNX(nodes) nodes);                                              \
  R *PNX(get_f_real)(const PNX(nodes) nodes);                                                 \
  R *PNX(get_grad_f_real)(const PNX(nodes) nodes);                                            \
  R *PNX(get_hessian_f_real)(const PNX(nodes) nodes);                                         \
  R *PNX(get_x)(const PNX(nodes) nodes);                                                      \
  C *PNX(get_f_hat)(const PNX(plan) ths);                                                     \
  R *PNX(get_f_hat_real)(const PNX(plan) ths);                                                \
  int PNX(get_d)(const PNX(plan) ths);                                                        \
  int PNX(get_m)(const PNX(plan) ths);                                                        \
  void PNX(get_x_max)(const PNX(plan) ths, R *x_max);                                         \
  void PNX(get_N)(const PNX(plan) ths, INT *N);                                               \
  void PNX(get_n)(const PNX(plan) ths, INT *n);                                               \
  unsigned PNX(get_pnfft_flags)(const PNX(plan) ths);                                         \
  unsigned PNX(get_pfft_flags)(const PNX(plan) ths);                                          \
  void PNX(get_b)(const PNX(plan) ths, R *b0, R *b1, R *b2);                                  \
  /* api/pnfft.h:183-189; api/api-basic.c:199-244, :344-378, :450-455 */                      \
  void PNX(finalize)(PNX(plan) ths, unsigned pnfft_finalize_flags);                           \
  void PNX(trafo)(PNX(plan) ths, PNX(nodes) nodes, unsigned compute_flags);                   \
  void PNX(adj)(PNX(plan) ths, PNX(nodes) nodes, unsigned compute_flags);                     \
  /* api/pnfft.h:191-199; api/api-basic.c:27-34, kernel/malloc.c:25-55 */                     \
  void PNX(init)(void);                                                                       \
  void PNX(cleanup)(void);                                                                    \
  void *PNX(malloc)(size_t n);                                                                \
  R *PNX(alloc_real)(size_t n);                                                               \
  C *PNX(alloc_complex)(size_t n);                                                            \
  void PNX(free)(void *p);                                                                    \
  /* api/pnfft.h:201-216; api/api-basic.c:335-341, :663-704, api/api-adv.c:35-86 */           \
  void PNX(init_f_hat_3d)(const INT *N, const INT *local_N, const INT *local_N_start,          \
      unsigned pnfft_flags, C *data);                                                         \
  void PNX(init_f)(INT local_M, C *data);                                                     \
  void PNX(init_x_3d)(const R *lo, const R *up, INT loc_M, R *x);                             \
  void PNX(init_x_3d_adv)(const R *lo, const R *up, const R *x_max, INT loc_M, R *x);         \
  void PNX(zero_f_hat)(PNX(plan) ths);                                                        \
  /* api/pnfft.h:218-227; kernel/matrix_D.c:191-225, kernel/ndft-parallel.c:2288-2334 */      \
  R PNX(inv_phi_hat)(const PNX(plan) ths, int dim, INT k);                                    \
  R PNX(phi_hat)(const PNX(plan) ths, int dim, INT k);                                        \
  R PNX(psi)(const PNX(plan) ths, int dim, R x);                                              \
  R PNX(dpsi)(const PNX(plan) ths, int dim, R x);                                             \
  R PNX(ddpsi)(const PNX(plan) ths, int dim, R x);                                            \
  /* api/pnfft.h:240-264; kernel/timer.c:43-361 (stage times from HIP events) */              \
  double *PNX(get_timer_trafo)(PNX(plan) ths);                                                \
  double *PNX(get_timer_adj)(PNX(plan) ths);                                                  \
  void PNX(timer_average)(double *timer);                                                     \
  double *PNX(timer_copy)(const double *orig);                                                \
  double *PNX(timer_reduce_max)(MPI_Comm comm, double *timer);                                \
  double *PNX(timer_add)(const double *sum1, const double *sum2);                             \
  void PNX(timer_free)(double *ths);                                                          \
  void PNX(reset_timer)(PNX(plan) ths);                                                       \
  void PNX(print_average_timer)(const PNX(plan) ths, MPI_Comm comm);                          \
  void PNX(print_average_timer_adv)(const PNX(plan) ths, MPI_Comm comm);                      \
  /* Extensions (no reference counterpart).                                                   \
   * malloc_device: device memory freed by PNX(free); such pointers may be passed             \
   * to the set_* functions for device-resident execution.                                    \
   * init_nodes_device: like init_nodes, but the malloc'd node arrays live in                 \
   * device memory (get_x etc. return device pointers).                                       \
   * trafo_grad_x_vjp: for w = cotangent of f (C[M]) writes                                   \
   *   grad_x[3j+t] = Re(conj(w_j) * d f_j / d x_t)  (R[3M]),                                 \
   * the x-adjoint of f = trafo(f_hat), computed by the window-derivative                     \
   * gather fused with the contraction (autograd backward for x).                             \
   * node_bins: bins the nodes for the plan's grid (if not binned yet) and copies              \
   * the node processing order (node ids sorted stably by bin key, int[M]) and                \
   * the sorted keys (unsigned[M]) to host or device arrays; returns 0 on success.            \
   * A node's key is tile * T0 T1 T2 + cell: per axis its stencil start                        \
   * s_t = (floor(n_t x_t) - m) mod n_t gives tile index s_t / T_t (tiles                      \
   * row-major, axis 2 fastest) and offset l_t = s_t mod T_t; with q = min(T2, 4)              \
   * the cell is ((l2 / q) T0 T1 + l0 T1 + l1) q + l2 mod q.  T_t is the plan's                \
   * tile size (8 for m <= 4 and for FP64 m <= 6, else 4; halved where it                     \
   * does not divide n_t).                                                                     \
   * kernel_times: per-kernel device time (ms, from HIP events on the plan's                   \
   * stream) and launch counts accumulated since the plan was made or                         \
   * reset_timer was called, for the kernel classes PNFFT_KT_* below; copies                  \
   * min(n, PNFFT_KT_COUNT) entries and returns PNFFT_KT_COUNT.                               \
   * oversampled_grid: device pointer to the plan's oversampled grid, n0*n1*n2                \
   * complex values (reals for c2r plans) in natural DFT order (index l_t mod                 \
   * n_t, row-major); after trafo it holds F D f_hat (F f_hat with                            \
   * PNFFT_OMIT_DECONV), after adj B^T f.                                                     \
   * The plan's D factors and window values carry reciprocal powers of two per               \
   * axis (2^e and 2^-e, e = exponent of psi(0)), so the grid is 2^e times the               \
   * reference's g2 (exact scaling; B F D and, with PNFFT_OMIT_DECONV, B F are                \
   * unchanged).                                                                              \
   * last_error: why the most recent plan creation, trafo, adj, precompute_psi,               \
   * trafo_grad_x_vjp or node binning call failed, or NULL if it succeeded.                   \
   * node_bins_noalloc: no counterpart in the reference; for callers that cannot               \
   * allocate device memory (code under HIP graph capture, tight serving loops).              \
   * Like node_bins for M nodes whose coordinates x (R[3M]) and outputs order                 \
   * (int[M]) and keys (unsigned[M]) are device arrays of the caller; allocates               \
   * nothing: the sort's scratch is a static device array of the library                      \
   * (about 17 MB per GPU, held until the process ends) for at most 2^20 nodes;              \
   * larger M is rejected (-1, last_error set).  Calls that overlap on one GPU                \
   * share that array and race.  Returns 0 on success. */                                     \
  void *PNX(malloc_device)(size_t n);                                                         \
  PNX(nodes) PNX(init_nodes_device)(INT local_M, unsigned malloc_flags);                      \
  void PNX(trafo_grad_x_vjp)(PNX(plan) ths, PNX(nodes) nodes, const C *w, R *grad_x);         \
  int PNX(node_bins)(PNX(plan) ths, PNX(nodes) nodes, int *order, unsigned *keys);            \
  int PNX(kernel_times)(PNX(plan) ths, double *ms, long *launches, int n);                     \
  int PNX(node_bins_noalloc)(PNX(plan) ths, const R *x, INT M, int *order, unsigned *keys);   \
  C *PNX(oversampled_grid)(PNX(plan) ths);                                                    \
  const char *PNX(last_error)(void);

/* kernel classes of PNX(kernel_times) (extension) */
#define PNFFT_KT_FFT_ROWS 0        /* axis-2 FFT pass (D or D^T fused) */
#define PNFFT_KT_FFT_COLS1 1       /* axis-1 FFT pass */
#define PNFFT_KT_FFT_COLS0 2       /* axis-0 FFT pass */
#define PNFFT_KT_GATHER_TILED 3    /* trafo B, LDS-tiled kernel */
#define PNFFT_KT_GATHER_DIRECT 4   /* trafo B, direct kernel (sparse tiles) */
#define PNFFT_KT_SCATTER_TILED 5   /* adj B^T, LDS-tiled kernel */
#define PNFFT_KT_SCATTER_DIRECT 6  /* adj B^T, direct kernel */
#define PNFFT_KT_BINNING 7         /* node binning (keys, radix sort, work items) */
#define PNFFT_KT_SCATTER_FOLD 8    /* adj B^T, sum of the tiles' subgrids into the grid */
#define PNFFT_KT_COUNT 9

#define PNFFT_MANGLE_DOUBLE(name) pnfft_##name
#define PNFFT_MANGLE_FLOAT(name) pnfftf_##name

PNFFT_DECLARE_API(PNFFT_MANGLE_DOUBLE, double, pnfft_complex, ptrdiff_t)
PNFFT_DECLARE_API(PNFFT_MANGLE_FLOAT, float, pnfftf_complex, ptrdiff_t)

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifndef PNFFT_PI
#define PNFFT_PI 3.14159265358979323846
#endif

/* Flags for plan initialization: values of api/pnfft.h:302-334 (aliasing kept). */
#define PNFFT_PRE_PHI_HAT           (1U << 0)
#define PNFFT_PRE_PHI_HUT           ((PNFFT_PRE_PHI_HAT))
#define PNFFT_FAST_GAUSSIAN         (1U << 1)
#define PNFFT_PRE_CONST_PSI         (1U << 2)
#define PNFFT_PRE_LIN_PSI           (1U << 3)
#define PNFFT_PRE_QUAD_PSI          (1U << 4)
#define PNFFT_PRE_CUB_PSI           (1U << 5)
#define PNFFT_PRE_INTPOL_PSI        ((PNFFT_PRE_CONST_PSI | PNFFT_PRE_LIN_PSI | PNFFT_PRE_QUAD_PSI | PNFFT_PRE_CUB_PSI))
#define PNFFT_MALLOC_F_HAT          (1U << 6)
#define PNFFT_FFT_OUT_OF_PLACE      (0U)
#define PNFFT_FFT_IN_PLACE          (1U << 7)
#define PNFFT_INTERLACED            (1U << 8)
#define PNFFT_SHIFTED_F_HAT         (1U << 9)
#define PNFFT_SHIFTED_X             (1U << 10)
#define PNFFT_TRANSPOSED_NONE       (0U)
#define PNFFT_TRANSPOSED_F_HAT      (1U << 11)
#define PNFFT_DIFF_AD               (0U)
#define PNFFT_DIFF_IK               (1U << 12)
#define PNFFT_WINDOW_KAISER_BESSEL  (0U)
#define PNFFT_WINDOW_GAUSSIAN       (1U << 13)
#define PNFFT_WINDOW_BSPLINE        (1U << 14)
#define PNFFT_WINDOW_SINC_POWER     (1U << 15)
#define PNFFT_WINDOW_BESSEL_I0      (1U << 16)
#define PNFFT_USE_FK_GAUSSIAN_T     (1U << 17)
#define PNFFT_WINDOW_GAUSSIAN_T     ((PNFFT_USE_FK_GAUSSIAN_T | PNFFT_WINDOW_GAUSSIAN))
#define PNFFT_SORT_NODES            (1U << 18)

/* Plan finalization (api/pnfft.h:340). */
#define PNFFT_FREE_F_HAT            ((PNFFT_MALLOC_F_HAT))

/* Node initialization (api/pnfft.h:345-353). */
#define PNFFT_MALLOC_NONE           (0U)
#define PNFFT_MALLOC_X              (1U << 0)
#define PNFFT_MALLOC_F              (1U << 1)
#define PNFFT_MALLOC_GRAD_F         (1U << 2)
#define PNFFT_MALLOC_HESSIAN_F      (1U << 3)
#define PNFFT_MALLOC_ALL            ((PNFFT_MALLOC_X | PNFFT_MALLOC_F | PNFFT_MALLOC_GRAD_F | PNFFT_MALLOC_HESSIAN_F))
#define PNFFT_REAL_F                (1U << 4)

/* Window precomputation (api/pnfft.h:359-364). */
#define PNFFT_PRE_TENSOR            (0U)
#define PNFFT_PRE_FULL              (1U << 0)
#define PNFFT_PRE_PSI               (1U << 1)
#define PNFFT_PRE_GRAD_PSI          (1U << 2)
#define PNFFT_PRE_HESSIAN_PSI       (1U << 3)

/* Node finalization (api/pnfft.h:372-377). */
#define PNFFT_FREE_NONE             (0U)
#define PNFFT_FREE_X                ((PNFFT_MALLOC_X))
#define PNFFT_FREE_F                ((PNFFT_MALLOC_F))
#define PNFFT_FREE_GRAD_F           ((PNFFT_MALLOC_GRAD_F))
#define PNFFT_FREE_HESSIAN_F        ((PNFFT_MALLOC_HESSIAN_F))
#define PNFFT_FREE_ALL              ((PNFFT_MALLOC_ALL))

/* Execution (api/pnfft.h:383-390). */
#define PNFFT_COMPUTE_F             (1U << 0)
#define PNFFT_COMPUTE_GRAD_F        (1U << 1)
#define PNFFT_COMPUTE_HESSIAN_F     (1U << 2)
#define PNFFT_COMPUTE_DIRECT        (1U << 3)
#define PNFFT_COMPUTE_ACCUMULATED   (1U << 4)
#define PNFFT_OMIT_DECONV           (1U << 5)
#define PNFFT_OMIT_FFT              (1U << 6)
#define PNFFT_OMIT_CONV             (1U << 7)

/* Timer slots (api/pnfft.h:407-418). */
#define PNFFT_TIMER_ITER            (0)
#define PNFFT_TIMER_WHOLE           (1)
#define PNFFT_TIMER_LOOP_B          (2)
#define PNFFT_TIMER_SORT_NODES      (3)
#define PNFFT_TIMER_GCELLS          (4)
#define PNFFT_TIMER_MATRIX_B        (5)
#define PNFFT_TIMER_MATRIX_F        (6)
#define PNFFT_TIMER_MATRIX_D        (7)
#define PNFFT_TIMER_SHIFT_INPUT     (8)
#define PNFFT_TIMER_SHIFT_OUTPUT    (9)
#define PNFFT_TIMER_LENGTH          (10)

#endif /* PNFFT_H */
