"""MI355X (gfx950) single-GPU NFFT with PNFFT's C plan interface.

The product is ``libpnfft_amd.so`` (HIP kernels + the C interface declared in
``include/pnfft.h``).  This package loads it through ctypes and mirrors the C
interface one-to-one (``lib().d`` for the ``pnfft_`` double functions,
``lib().f`` for the ``pnfftf_`` single functions).  There is no Python or CPU
fallback: without the built library every entry point raises.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpnfft_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pnfft.h")

# flag values of include/pnfft.h (api/pnfft.h:302-390)
PNFFT_PRE_PHI_HAT = 1 << 0
PNFFT_FAST_GAUSSIAN = 1 << 1
PNFFT_MALLOC_F_HAT = 1 << 6
PNFFT_INTERLACED = 1 << 8
PNFFT_DIFF_IK = 1 << 12
PNFFT_WINDOW_KAISER_BESSEL = 0
PNFFT_WINDOW_GAUSSIAN = 1 << 13
PNFFT_WINDOW_BSPLINE = 1 << 14
PNFFT_WINDOW_SINC_POWER = 1 << 15
PNFFT_WINDOW_BESSEL_I0 = 1 << 16
PNFFT_SORT_NODES = 1 << 18
PNFFT_REAL_F = 1 << 4
PNFFT_FREE_F_HAT = PNFFT_MALLOC_F_HAT
PNFFT_MALLOC_X = 1 << 0
PNFFT_MALLOC_F = 1 << 1
PNFFT_MALLOC_GRAD_F = 1 << 2
PNFFT_FREE_X = PNFFT_MALLOC_X
PNFFT_FREE_F = PNFFT_MALLOC_F
PNFFT_FREE_GRAD_F = PNFFT_MALLOC_GRAD_F
PNFFT_FREE_ALL = 0xF
PNFFT_PRE_TENSOR = 0
PNFFT_PRE_FULL = 1 << 0
PNFFT_PRE_PSI = 1 << 1
PNFFT_PRE_GRAD_PSI = 1 << 2
PNFFT_PRE_HESSIAN_PSI = 1 << 3
PNFFT_COMPUTE_F = 1 << 0
PNFFT_COMPUTE_GRAD_F = 1 << 1
PNFFT_COMPUTE_HESSIAN_F = 1 << 2
PNFFT_COMPUTE_DIRECT = 1 << 3
PNFFT_COMPUTE_ACCUMULATED = 1 << 4
PNFFT_OMIT_DECONV = 1 << 5
PNFFT_OMIT_FFT = 1 << 6
PNFFT_OMIT_CONV = 1 << 7
PNFFT_TIMER_ITER = 0
PNFFT_TIMER_WHOLE = 1
PNFFT_TIMER_LOOP_B = 2
PNFFT_TIMER_SORT_NODES = 3
PNFFT_TIMER_MATRIX_B = 5
PNFFT_TIMER_MATRIX_F = 6
PNFFT_TIMER_MATRIX_D = 7
MPI_COMM_WORLD = 0x44000000


class LibraryMissing(RuntimeError):
    pass


_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint
_PD = ctypes.c_ssize_t
_PDP = ctypes.POINTER(ctypes.c_ssize_t)
_DBLP = ctypes.POINTER(ctypes.c_double)


class _Api:
    """ctypes view of one precision of the C interface (pnfft_ or pnfftf_)."""

    def __init__(self, cdll, prefix, real):
        self.prefix = prefix
        self.real = real
        RP = ctypes.POINTER(real)

        def fn(name, res, *args):
            f = getattr(cdll, prefix + name)
            f.restype = res
            f.argtypes = list(args)
            setattr(self, name, f)

        fn("init_guru", _P, _I, _PDP, _PDP, RP, _I, _U, _U, _I)
        fn("init_guru_c2r", _P, _I, _PDP, _PDP, RP, _I, _U, _U, _I)
        fn("init_adv", _P, _I, _PDP, _U, _U, _I)
        fn("init_3d", _P, _PDP, _I)
        fn("init_3d_c2r", _P, _PDP, _I)
        fn("local_size_3d", None, _PDP, _I, _U, _PDP, _PDP, RP, RP)
        fn("local_size_guru", None, _I, _PDP, _PDP, RP, _I, _I, _U, _PDP, _PDP, RP, RP)
        fn("init_nodes", _P, _PD, _U)
        fn("init_nodes_device", _P, _PD, _U)
        fn("free_nodes", None, _P, _U)
        fn("precompute_psi", None, _P, _P, _U)
        fn("precompute_one_psi", None, _P, _P, _U)
        for nm in ("set_f", "set_grad_f", "set_hessian_f", "set_f_real", "set_grad_f_real", "set_x"):
            fn(nm, None, _P, _P)
        fn("set_f_hat", None, _P, _P)
        fn("set_f_hat_real", None, _P, _P)
        fn("set_b", None, real, real, real, _P)
        for nm in ("get_f", "get_grad_f", "get_hessian_f", "get_f_real", "get_grad_f_real", "get_x"):
            fn(nm, _P, _P)
        fn("get_f_hat", _P, _P)
        fn("get_f_hat_real", _P, _P)
        fn("get_d", _I, _P)
        fn("get_m", _I, _P)
        fn("get_x_max", None, _P, RP)
        fn("get_N", None, _P, _PDP)
        fn("get_n", None, _P, _PDP)
        fn("get_pnfft_flags", _U, _P)
        fn("get_pfft_flags", _U, _P)
        fn("get_b", None, _P, RP, RP, RP)
        fn("finalize", None, _P, _U)
        fn("trafo", None, _P, _P, _U)
        fn("adj", None, _P, _P, _U)
        fn("trafo_grad_x_vjp", None, _P, _P, _P, _P)
        fn("init", None)
        fn("cleanup", None)
        fn("malloc", _P, ctypes.c_size_t)
        fn("malloc_device", _P, ctypes.c_size_t)
        fn("alloc_real", _P, ctypes.c_size_t)
        fn("alloc_complex", _P, ctypes.c_size_t)
        fn("free", None, _P)
        fn("inv_phi_hat", real, _P, _I, _PD)
        fn("phi_hat", real, _P, _I, _PD)
        fn("psi", real, _P, _I, real)
        fn("dpsi", real, _P, _I, real)
        fn("ddpsi", real, _P, _I, real)
        fn("get_timer_trafo", _DBLP, _P)
        fn("get_timer_adj", _DBLP, _P)
        fn("timer_free", None, _DBLP)
        fn("reset_timer", None, _P)
        fn("node_bins", _I, _P, _P, _P, _P)
        fn("node_bins_noalloc", _I, _P, _P, _PD, _P, _P)
        fn("kernel_times", _I, _P, _P, _P, _I)
        fn("oversampled_grid", _P, _P)
        fn("last_error", ctypes.c_char_p)

    # convenience wrappers -------------------------------------------------
    def arr3(self, v, ctype=None):
        ctype = ctype or _PD
        return (ctype * 3)(*v)


class _Lib:
    def __init__(self, cdll):
        self.cdll = cdll
        self.d = _Api(cdll, "pnfft_", ctypes.c_double)
        self.f = _Api(cdll, "pnfftf_", ctypes.c_float)

    def api(self, precision):
        return self.d if precision in ("double", "f64", 8) else self.f


_lib = None


def lib():
    """Load libpnfft_amd.so (raises LibraryMissing if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is missing: run pnfft_amd/build.py (or __graft_entry__.build())")
        _lib = _Lib(ctypes.CDLL(LIB_PATH))
    return _lib


def header_symbols():
    """Every function name declared by include/pnfft.h (both precisions)."""
    txt = open(HEADER).read()
    body = txt[txt.index("#define PNFFT_DECLARE_API"): txt.index("#define PNFFT_MANGLE_DOUBLE")]
    names = sorted(set(re.findall(r"PNX\((\w+)\)\s*\(", body)))
    return [p + n for p in ("pnfft_", "pnfftf_") for n in names]
