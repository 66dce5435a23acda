"""FP64 CPU oracle for PNFFT's single-process NFFT path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py`` -- never by the product package.

The C part (``pnfft_oracle.c``) restates the reference's windows, matrix D,
matrix B (on the ghost-padded grid) and the NDFT; this module composes them
into ``trafo``/``adj`` the way ``api/api-basic.c:170-196`` and ``:307-333`` do,
with the FFT stage (PFFT in the reference, ``kernel/ndft-parallel.c:1524-1546``)
taken by scipy.fft (pocketfft) on the host.
"""
import ctypes
import os

import numpy as np
import scipy.fft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# flag values of api/pnfft.h:302-390
PNFFT_FAST_GAUSSIAN = 1 << 1
PNFFT_WINDOW_KAISER_BESSEL = 0
PNFFT_WINDOW_GAUSSIAN = 1 << 13
PNFFT_WINDOW_BSPLINE = 1 << 14
PNFFT_WINDOW_SINC_POWER = 1 << 15
PNFFT_WINDOW_BESSEL_I0 = 1 << 16
PNFFT_PRE_CONST_PSI = 1 << 2
PNFFT_PRE_LIN_PSI = 1 << 3
PNFFT_PRE_QUAD_PSI = 1 << 4
PNFFT_PRE_CUB_PSI = 1 << 5
PNFFT_USE_FK_GAUSSIAN_T = 1 << 17
PNFFT_INTERLACED = 1 << 8
PNFFT_DIFF_IK = 1 << 12
PNFFT_COMPUTE_F = 1 << 0
PNFFT_COMPUTE_GRAD_F = 1 << 1
PNFFT_COMPUTE_HESSIAN_F = 1 << 2
PNFFT_COMPUTE_ACCUMULATED = 1 << 4
# Hessian component order (kernel/assign.c:905-921): xx, xy, xz, yy, yz, zz
HESS_PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))

W_KB, W_GAUSS, W_BSPLINE, W_SINCPOW, W_BESSEL_I0 = range(5)

_lib = None
_lib_fast = None
_FAST_PATH = os.path.join(_HERE, "_fast", "liboracle_fast.so")


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib_fast():
    """The same C restatement built on this host with the reference's release flags
    (-O3 -march=native -ffast-math, scripts/install_pnfft_gcc.sh:42): CPU-baseline timing only."""
    global _lib_fast
    if _lib_fast is None:
        import subprocess
        subprocess.run(["make", "-s", "-C", _HERE, "fast"], check=True)
        _lib_fast = _load(_FAST_PATH)
    return _lib_fast


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


def _load(path):
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lp = ctypes.POINTER(ctypes.c_long)
    L.orc_window_id.argtypes = [ctypes.c_uint]
    L.orc_window_id.restype = ctypes.c_int
    L.orc_window_param_b.argtypes = [lp, lp, ctypes.c_int, ctypes.c_uint, dp]
    L.orc_bessel_i0.argtypes = [ctypes.c_double]
    L.orc_bessel_i0.restype = ctypes.c_double
    L.orc_bessel_i1.argtypes = [ctypes.c_double]
    L.orc_bessel_i1.restype = ctypes.c_double
    L.orc_sinc.argtypes = [ctypes.c_double]
    L.orc_sinc.restype = ctypes.c_double
    L.orc_bspline.argtypes = [ctypes.c_int, ctypes.c_double]
    L.orc_bspline.restype = ctypes.c_double
    L.orc_inv_phi_hat.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_double, ctypes.c_int]
    L.orc_inv_phi_hat.restype = ctypes.c_double
    L.orc_psi_axis.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                               ctypes.c_double, dp, dp]
    L.orc_psi.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.orc_psi.restype = ctypes.c_double
    L.orc_dpsi.argtypes = L.orc_psi.argtypes
    L.orc_dpsi.restype = ctypes.c_double
    L.orc_ddpsi.argtypes = L.orc_psi.argtypes
    L.orc_ddpsi.restype = ctypes.c_double
    L.orc_psi_axis2.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                                ctypes.c_double, dp, dp, dp]
    L.orc_B_trafo_hess.argtypes = [lp, lp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_long, dp, dp,
                                   ctypes.c_int, dp]
    L.orc_ndft_hessian_c2c.argtypes = [lp, ctypes.c_long, dp, dp, dp]
    L.orc_cerf.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
    L.orc_intpol_table.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_long, ctypes.c_int, dp]
    L.orc_intpol_set.argtypes = [ctypes.c_int, ctypes.c_long] + [dp] * 9
    L.orc_intpol_clear.argtypes = []
    L.orc_intpol_axis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
    L.orc_ndft_hessian_c2r.argtypes = [lp, ctypes.c_long, dp, dp, dp]
    L.orc_ndft_trafo_c2c.argtypes = [lp, ctypes.c_long, dp, dp, dp, dp, ctypes.c_uint]
    L.orc_ndft_adj_c2c.argtypes = [lp, ctypes.c_long, dp, dp, dp, dp, ctypes.c_uint]
    L.orc_ndft_trafo_c2r.argtypes = [lp, ctypes.c_long, dp, dp, dp, dp, ctypes.c_uint]
    L.orc_inv_phi_hat_tables.argtypes = [lp, lp, ctypes.c_int, ctypes.c_uint, dp]
    L.orc_D.argtypes = [lp, dp, dp, dp, ctypes.c_int]
    L.orc_B_trafo.argtypes = [lp, lp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_long, dp, dp,
                              ctypes.c_int, dp, dp, ctypes.c_uint]
    L.orc_B_adj.argtypes = [lp, lp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_long, dp, dp, dp,
                            ctypes.c_int, dp, ctypes.c_uint]
    L.orc_B_adj_blocks.argtypes = L.orc_B_adj.argtypes
    L.orc_sort_nodes.argtypes = [lp, ctypes.c_int, ctypes.c_long, dp, lp]
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))


def _longs(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64).reshape(3))


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# plan-level parameters
# ---------------------------------------------------------------------------
def fft_output_size(n, x_max, m):
    """api/api-guru.c:195-218: no_t = min(n_t, 2*(floor(n_t x_max_t) + m + 2))."""
    return [min(int(nt), 2 * (int(np.floor(nt * xm)) + m + 2)) for nt, xm in zip(n, x_max)]


def window_id(flags):
    return lib().orc_window_id(flags)


def window_b(N, n, m, flags):
    b = np.zeros(3)
    lib().orc_window_param_b(_lp(_longs(N)), _lp(_longs(n)), m, flags, _dp(b))
    return b


def bessel_i0(x):
    return lib().orc_bessel_i0(float(x))


def bessel_i1(x):
    return lib().orc_bessel_i1(float(x))


def sinc(x):
    return lib().orc_sinc(float(x))


def bspline(k, x):
    return lib().orc_bspline(int(k), float(x))


def inv_phi_hat(window, k, n, b, m):
    return lib().orc_inv_phi_hat(window, int(k), int(n), float(b), int(m))


def inv_phi_hat_tables(N, n, m, flags):
    N = _longs(N)
    tab = np.zeros(int(N.sum()))
    lib().orc_inv_phi_hat_tables(_lp(N), _lp(_longs(n)), m, flags, _dp(tab))
    return tab


def psi_axis(window, n, b, m, nx, want_dpsi=True):
    c = 2 * m + 1
    psi = np.zeros(c)
    dpsi = np.zeros(c) if want_dpsi else None
    lib().orc_psi_axis(window, int(n), float(b), int(m), float(nx), _dp(psi), _dp(dpsi))
    return (psi, dpsi) if want_dpsi else psi


def psi(window, n, b, m, x):
    return lib().orc_psi(window, int(n), float(b), int(m), float(x))


def dpsi(window, n, b, m, x):
    return lib().orc_dpsi(window, int(n), float(b), int(m), float(x))


def cerf(z):
    """erf(z) for Re z > 0 (Abramowitz & Stegun 7.1.29; the reference calls libcerf, cerf/)."""
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().orc_cerf(float(z.real), float(z.imag), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def intpol_order(flags):
    """Interpolation order of kernel/ndft-parallel.c:997-1006 after the flag implications of
    api/api-guru.c:152-155 (PRE_HESSIAN_PSI -> PRE_GRAD_PSI -> PRE_PSI on the plan flags, whose
    values alias PRE_LIN_PSI -> PRE_CONST_PSI -> FAST_GAUSSIAN): -1 for none."""
    if flags & (1 << 3):
        flags |= 1 << 2
    if flags & (1 << 2):
        flags |= 1 << 1
    for bit, order in ((PNFFT_PRE_CONST_PSI, 0), (PNFFT_PRE_LIN_PSI, 1), (PNFFT_PRE_QUAD_PSI, 2),
                       (PNFFT_PRE_CUB_PSI, 3)):
        if flags & bit:
            return order
    return -1


def intpol_nodes(m):
    """Interpolation nodes per unit interval, kernel/ndft-parallel.c:1095: ceil(31/(2m+1)) * 2048."""
    return int(np.ceil(31.0 / (2 * m + 1))) * 2048


def intpol_table(window, n, b, m, order, deriv):
    """init_intpol_table_psi (kernel/ndft-parallel.c:318-352)."""
    K = intpol_nodes(m)
    tab = np.zeros(K * (2 * m + 1) * (order + 1))
    lib().orc_intpol_table(window, int(n), float(b), int(m), int(order), K, int(deriv), _dp(tab))
    return tab


def ddpsi(window, n, b, m, x):
    """Scalar psi''(x) (PNX(ddpsi), kernel/ndft-parallel.c:2320-2460)."""
    return lib().orc_ddpsi(window, int(n), float(b), int(m), float(x))


def ddpsi_axis(window, n, b, m, nx):
    """(psi, psi', psi'') on one axis's stencil (pre_ddpsi_tensor_*, kernel/ndft-parallel.c:1956-2105)."""
    c = 2 * m + 1
    out = [np.zeros(c) for _ in range(3)]
    lib().orc_psi_axis2(window, int(n), float(b), int(m), float(nx), *[_dp(a) for a in out])
    return tuple(out)


# ---------------------------------------------------------------------------
# NDFT
# ---------------------------------------------------------------------------
def ndft_trafo(N, x, fhat, compute_grad=False, compute_f=True):
    """Direct sums of kernel/ndft-parallel.c:377-613 (c2c). Returns (f, grad)."""
    N = _longs(N)
    x = _f64(x).reshape(-1)
    M = x.size // 3
    fhat = _c128(fhat).reshape(-1)
    assert fhat.size == int(np.prod(N))
    f = np.zeros(M, np.complex128)
    grad = np.zeros(3 * M, np.complex128) if compute_grad else None
    fl = (PNFFT_COMPUTE_F if compute_f else 0) | (PNFFT_COMPUTE_GRAD_F if compute_grad else 0)
    lib().orc_ndft_trafo_c2c(_lp(N), M, _dp(x), _dp(fhat.view(np.float64)), _dp(f.view(np.float64)),
                             _dp(grad.view(np.float64)) if grad is not None else None, fl)
    return (f if compute_f else None), grad


def ndft_hessian(N, x, fhat):
    """NDFT Hessian, c2c (kernel/ndft-parallel.c:450-510, :604-612): C[6M] in HESS_PAIRS order."""
    N = _longs(N)
    x = _f64(x).reshape(-1)
    M = x.size // 3
    fhat = _c128(fhat).reshape(-1)
    h = np.zeros(6 * M, np.complex128)
    lib().orc_ndft_hessian_c2c(_lp(N), M, _dp(x), _dp(fhat.view(np.float64)), _dp(h.view(np.float64)))
    return h


def ndft_hessian_c2r(N, x, fhat_half):
    """NDFT Hessian of a c2r plan (kernel/ndft-parallel.c:465-476, :604-609): R[6M]."""
    N = _longs(N)
    x = _f64(x).reshape(-1)
    M = x.size // 3
    fh = _c128(fhat_half).reshape(-1)
    h = np.zeros(6 * M)
    lib().orc_ndft_hessian_c2r(_lp(N), M, _dp(x), _dp(fh.view(np.float64)), _dp(h))
    return h


def ndft_adj(N, x, f=None, grad=None, fhat_in=None):
    """Direct adjoint sums of kernel/ndft-parallel.c:617-722 (adds onto fhat_in)."""
    N = _longs(N)
    x = _f64(x).reshape(-1)
    M = x.size // 3
    fhat = np.zeros(int(np.prod(N)), np.complex128) if fhat_in is None else _c128(fhat_in).reshape(-1).copy()
    fl = 0
    fa = ga = None
    if f is not None:
        fa = _c128(f).reshape(-1)
        fl |= PNFFT_COMPUTE_F
    if grad is not None:
        ga = _c128(grad).reshape(-1)
        fl |= PNFFT_COMPUTE_GRAD_F
    lib().orc_ndft_adj_c2c(_lp(N), M, _dp(x), _dp(fa.view(np.float64)) if fa is not None else None,
                           _dp(ga.view(np.float64)) if ga is not None else None,
                           _dp(fhat.view(np.float64)), fl)
    return fhat


def ndft_trafo_c2r(N, x, fhat_half, compute_grad=False):
    N = _longs(N)
    x = _f64(x).reshape(-1)
    M = x.size // 3
    fh = _c128(fhat_half).reshape(-1)
    assert fh.size == int(N[0] * N[1] * (N[2] // 2 + 1))
    f = np.zeros(M)
    grad = np.zeros(3 * M) if compute_grad else None
    fl = PNFFT_COMPUTE_F | (PNFFT_COMPUTE_GRAD_F if compute_grad else 0)
    lib().orc_ndft_trafo_c2r(_lp(N), M, _dp(x), _dp(fh.view(np.float64)), _dp(f), _dp(grad), fl)
    return f, grad


def ndft_adj_c2r(N, x, f=None, grad=None):
    """Adjoint NDFT of a c2r plan (kernel/ndft-parallel.c:617-722 over the half block k2 in [-N2/2, 0])
    for real f / grad: the c2c adjoint sums restricted to the stored half spectrum."""
    full = ndft_adj(N, x, f=None if f is None else np.asarray(f, np.complex128),
                    grad=None if grad is None else np.asarray(grad, np.complex128))
    return full.reshape(N)[:, :, : N[2] // 2 + 1].reshape(-1)


def semirandom_hermitian(N, half=False):
    """Deterministic Hermitian input of tests/check_trafo_vs_ndft_c2r.c:202-231:
    semirandom(k) + conj(semirandom(-k)), zero wherever a component equals -N_t/2."""
    def D(i):
        return 1000.0 / (1 if i == 0 else i)

    def sr(k0, k1, k2):
        if k0 == -(N[0] // 2) or k1 == -(N[1] // 2) or k2 == -(N[2] // 2):  # C's -N/2 truncates
            return 0j
        if k0 == (N[0] + 1) // 2:
            return sr(-k0, k1, k2)
        if k1 == (N[1] + 1) // 2:
            return sr(k0, -k1, k2)
        if k2 == (N[2] + 1) // 2:
            return sr(k0, k1, -k2)
        l0 = k0 + N[0] if k0 else 1
        l1 = k1 - 2 * N[1] if k1 else 1
        l2 = k2 + N[2] if k2 else 1
        re = D(l0 + l1 * l2) + 3 * D(l0 * l1 + l2) + D(l2 * l0 + l1)
        im = 3 * D(l0 + l1 * l1) + D(l1 + l2 * l2) - D(l2 + l0 * l0)
        return re + 1j * im

    n2 = N[2] // 2 + 1 if half else N[2]
    out = np.zeros((N[0], N[1], n2), np.complex128)
    for i0 in range(N[0]):
        for i1 in range(N[1]):
            for i2 in range(n2):
                k0, k1, k2 = i0 - N[0] // 2, i1 - N[1] // 2, i2 - N[2] // 2
                out[i0, i1, i2] = sr(k0, k1, k2) + np.conj(sr(-k0, -k1, -k2))
    return out


# ---------------------------------------------------------------------------
# Fast NFFT (D -> F -> ghost cells -> B), c2c
# ---------------------------------------------------------------------------
def _embed(B, n):
    """Centered array B (index k + N_t/2, k in [-N_t/2, N_t/2)) placed into a zero
    array of shape n in DFT order (k mod n_t); even N_t <= n_t."""
    if any(v % 2 for v in B.shape):
        A = np.zeros(n, B.dtype)
        idx = [np.arange(-(v // 2), v - v // 2) % nt for v, nt in zip(B.shape, n)]
        A[np.ix_(*idx)] = B
        return A
    B = np.fft.ifftshift(B)
    for t in range(3):
        Nt, nt = B.shape[t], n[t]
        if nt == Nt:
            continue
        h = Nt // 2
        zshape = list(B.shape)
        zshape[t] = nt - Nt
        lo, hi = np.split(B, [h], axis=t)
        B = np.concatenate([lo, np.zeros(zshape, B.dtype), hi], axis=t)
    return B


def _extract(G, N):
    """Inverse of _embed: the entries k in [-N_t/2, N_t/2) of the DFT-ordered G, centered."""
    if any(v % 2 for v in N):
        idx = [np.arange(-(v // 2), v - v // 2) % nt for v, nt in zip(N, G.shape)]
        return G[np.ix_(*idx)]
    for t in range(3):
        Nt, nt = N[t], G.shape[t]
        if nt == Nt:
            continue
        h = Nt // 2
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[t] = slice(0, h)
        hi[t] = slice(nt - h, nt)
        G = np.concatenate([G[tuple(lo)], G[tuple(hi)]], axis=t)
    return np.fft.fftshift(G)


class Plan:
    """Parameters of one oracle plan (api/api-guru.c:110-164, kernel/ndft-parallel.c:869-1069).

    c2r=True restates the real-valued plans (PNX(init_guru_c2r)): f_hat is the half spectrum
    [k0][k1][k2] with k2 in [-N2/2, 0] (N2/2+1 entries), the oversampled grid is real, and x,
    f, grad are real (kernel/ndft-parallel.c:2809-2847, :2972-2990).
    """

    def __init__(self, N, n=None, m=6, flags=0, x_max=(0.5, 0.5, 0.5), workers=-1, c2r=False):
        self.N = [int(v) for v in N]
        self.n = [2 * v for v in self.N] if n is None else [int(v) for v in n]
        self.m = int(m)
        self.flags = int(flags)
        self.x_max = list(x_max)
        self.no = fft_output_size(self.n, self.x_max, self.m)
        self.window = window_id(self.flags)
        self.b = window_b(self.N, self.n, self.m, self.flags)
        self.tab = inv_phi_hat_tables(self.N, self.n, self.m, self.flags)
        self.workers = workers
        self.c2r = bool(c2r)
        self.Nh = [self.N[0], self.N[1], self.N[2] // 2 + 1]  # c2r half spectrum extent
        # interpolated windows (PNFFT_PRE_*_PSI): tables [derivative][axis], the derivatives
        # only without ik-differentiation (kernel/ndft-parallel.c:1097-1122)
        self.ip_order = intpol_order(self.flags)
        self.ip_tabs = None
        if self.ip_order >= 0:
            derivs = 1 if self.flags & PNFFT_DIFF_IK else 3
            self.ip_tabs = [[intpol_table(self.window, self.n[t], self.b[t], self.m, self.ip_order, d)
                             for t in range(3)] for d in range(derivs)]
            zero = np.zeros(1)
            while len(self.ip_tabs) < 3:
                self.ip_tabs.append([zero, zero, zero])

    def _ip_on(self):
        if self.ip_tabs is not None:
            lib().orc_intpol_set(self.ip_order, intpol_nodes(self.m),
                                 *[_dp(self.ip_tabs[d][t]) for d in range(3) for t in range(3)])

    def _ip_off(self):
        if self.ip_tabs is not None:
            lib().orc_intpol_clear()

    # -- c2r: D on the half spectrum, k2 = i2 - N2/2 <= 0 (kernel/matrix_D.c:319-356 over local_N)
    def _D_half(self, fhat_half, accumulate_into=None):
        N0, N1, N2 = self.N
        t0 = self.tab[:N0]
        t1 = self.tab[N0:N0 + N1]
        t2 = self.tab[N0 + N1:N0 + N1 + self.Nh[2]]
        fac = (t0[:, None, None] * t1[None, :, None]) * t2[None, None, :]
        return np.asarray(fhat_half).reshape(self.Nh) * fac

    def _F_c2r(self, g1_half):
        """Pruned shifted c2r: the real grid g_l = sum over the Hermitian extension of the half.

        With the input shifted so that the stored k2 in [-N2/2, 0] lies in the FFTW half
        [0, n2/2], FFTW's c2r adds every stored coefficient with its conjugate partner except
        the k2 = 0 plane, which lands on the Nyquist slot and contributes its real part once.
        """
        N, n, no = self.N, self.n, self.no
        A = np.zeros(n, np.complex128)
        i0 = np.arange(-N[0] // 2, N[0] // 2) % n[0]
        i1 = np.arange(-N[1] // 2, N[1] // 2) % n[1]
        k2 = np.arange(-N[2] // 2, 1)
        w = np.where(k2 < 0, 2.0, 1.0)
        A[np.ix_(i0, i1, k2 % n[2])] = g1_half * w[None, None, :]
        G = scipy.fft.fftn(A, workers=self.workers).real
        oidx = [np.arange(-nt // 2, nt // 2) % nft for nt, nft in zip(no, n)]
        return G[np.ix_(*oidx)]

    def _FH_r2c(self, g2_real):
        N, n, no = self.N, self.n, self.no
        A = np.zeros(n, np.complex128)
        oidx = [np.arange(-nt // 2, nt // 2) % nft for nt, nft in zip(no, n)]
        A[np.ix_(*oidx)] = g2_real
        G = scipy.fft.ifftn(A, workers=self.workers) * float(np.prod(n))
        i0 = np.arange(-N[0] // 2, N[0] // 2) % n[0]
        i1 = np.arange(-N[1] // 2, N[1] // 2) % n[1]
        k2 = np.arange(-N[2] // 2, 1) % n[2]
        return G[np.ix_(i0, i1, k2)]

    def trafo_c2r(self, x, fhat_half, compute_f=True, compute_grad=False):
        x = _f64(x).reshape(-1)
        M = x.size // 3
        g2 = self._F_c2r(self._D_half(fhat_half))
        gpad = np.ascontiguousarray(np.pad(g2, [(self.m, self.m)] * 3, mode="wrap"))
        f = np.zeros(M) if compute_f else None
        grad = np.zeros(3 * M) if compute_grad else None
        fl = (PNFFT_COMPUTE_F if compute_f else 0) | (PNFFT_COMPUTE_GRAD_F if compute_grad else 0)
        self._ip_on()
        lib().orc_B_trafo(_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M,
                          _dp(x), _dp(gpad), 1, _dp(f), _dp(grad), fl)
        self._ip_off()
        return f, grad

    def adj_c2r(self, x, f=None, grad=None, parallel=False):
        x = _f64(x).reshape(-1)
        M = x.size // 3
        ng = [v + 2 * self.m for v in self.no]
        gpad = np.zeros(ng)
        fl = 0
        fa = ga = None
        if f is not None:
            fa = _f64(f).reshape(-1)
            fl |= PNFFT_COMPUTE_F
        if grad is not None:
            ga = _f64(grad).reshape(-1)
            fl |= PNFFT_COMPUTE_GRAD_F
        scatter = lib().orc_B_adj_blocks if parallel else lib().orc_B_adj
        self._ip_on()
        scatter(_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M,
                _dp(x), _dp(fa), _dp(ga), 1, _dp(gpad), fl)
        self._ip_off()
        g2 = self.ghost_reduce(gpad)
        return self._D_half(self._FH_r2c(g2)).reshape(-1)

    # -- D (kernel/matrix_D.c:227-280)
    def D(self, fhat):
        out = np.empty(int(np.prod(self.N)), np.complex128)
        src = _c128(fhat).reshape(-1)
        lib().orc_D(_lp(_longs(self.N)), _dp(self.tab), _dp(src.view(np.float64)), _dp(out.view(np.float64)), 0)
        return out

    def DT_accumulate(self, g1, fhat):
        src = _c128(g1).reshape(-1)
        lib().orc_D(_lp(_longs(self.N)), _dp(self.tab), _dp(src.view(np.float64)), _dp(fhat.view(np.float64)), 1)

    # -- F: pruned, shifted, unnormalized forward FFT N -> no (doc/manual.tex:178-243)
    def F(self, g1):
        N, n, no = self.N, self.n, self.no
        G = scipy.fft.fftn(_embed(np.asarray(g1).reshape(N), n), workers=self.workers)
        return _extract(G, no)

    # -- F^H: backward FFT no -> N
    def FH(self, g2):
        N, n, no = self.N, self.n, self.no
        G = scipy.fft.ifftn(_embed(np.asarray(g2).reshape(no), n), workers=self.workers)
        G *= float(np.prod(n))
        return _extract(G, N).reshape(-1)

    # -- ghost cells (PX(exchange)/PX(reduce), kernel/ndft-parallel.c:2558,2679):
    # m planes below, m above, periodic over the FFT output of size no.
    def ghost_exchange(self, g2):
        m = self.m
        return np.pad(g2, [(m, m)] * 3, mode="wrap")

    def ghost_reduce(self, gpad):
        """Fold every padded index i of each axis back to (i - m) mod no, axis by axis.
        Works in place on gpad (overwritten) and returns the folded no-shaped array."""
        m = self.m

        def sl(t, a, b):
            r = [slice(None)] * 3
            r[t] = slice(a, b)
            return tuple(r)

        for t in range(3):
            nt = self.no[t]
            core = gpad[sl(t, m, m + nt)]
            core[sl(t, nt - m, nt)] += gpad[sl(t, 0, m)]  # i in [0, m) -> nt - m + i
            core[sl(t, 0, m)] += gpad[sl(t, m + nt, nt + 2 * m)]  # i in [m + nt, nt + 2m) -> i - m - nt
        return np.ascontiguousarray(gpad[m:m + self.no[0], m:m + self.no[1], m:m + self.no[2]])

    def trafo(self, x, fhat, compute_f=True, compute_grad=False):
        """pnfft_trafo (api/api-basic.c:199-244) on host arrays; returns (f, grad)."""
        x = _f64(x).reshape(-1)
        M = x.size // 3
        g2 = self.F(self.D(fhat))
        gpad = np.ascontiguousarray(self.ghost_exchange(g2))
        f = np.zeros(M, np.complex128) if compute_f else None
        grad = np.zeros(3 * M, np.complex128) if compute_grad else None
        fl = (PNFFT_COMPUTE_F if compute_f else 0) | (PNFFT_COMPUTE_GRAD_F if compute_grad else 0)
        self._ip_on()
        lib().orc_B_trafo(_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M,
                          _dp(x), _dp(gpad.view(np.float64)), 0,
                          _dp(f.view(np.float64)) if f is not None else None,
                          _dp(grad.view(np.float64)) if grad is not None else None, fl)
        self._ip_off()
        return f, grad

    def adj(self, x, f=None, grad=None, fhat_in=None, parallel=False):
        """pnfft_adj (api/api-basic.c:344-378); adds onto fhat_in if given (ACCUMULATED).
        parallel: scatter on all OpenMP threads (orc_B_adj_blocks; same sums, grouped order)."""
        x = _f64(x).reshape(-1)
        M = x.size // 3
        ng = [v + 2 * self.m for v in self.no]
        gpad = np.zeros(ng, np.complex128)
        fl = 0
        fa = ga = None
        if f is not None:
            fa = _c128(f).reshape(-1)
            fl |= PNFFT_COMPUTE_F
        if grad is not None:
            ga = _c128(grad).reshape(-1)
            fl |= PNFFT_COMPUTE_GRAD_F
        scatter = lib().orc_B_adj_blocks if parallel else lib().orc_B_adj
        self._ip_on()
        scatter(_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M,
                _dp(x), _dp(fa.view(np.float64)) if fa is not None else None,
                _dp(ga.view(np.float64)) if ga is not None else None, 0,
                _dp(gpad.view(np.float64)), fl)
        self._ip_off()
        g2 = self.ghost_reduce(gpad)
        g1 = self.FH(g2)
        fhat = np.zeros(int(np.prod(self.N)), np.complex128) if fhat_in is None else _c128(fhat_in).reshape(-1).copy()
        self.DT_accumulate(g1, fhat)
        return fhat


    # -----------------------------------------------------------------------
    # interlacing, ik-differentiation and Hessians (c2c and c2r plans)
    # -----------------------------------------------------------------------
    def _kvecs(self):
        """Frequencies of f_hat's axes (local_N_start = -N/2; the c2r half spectrum has k2 <= 0)."""
        ks = [np.arange(-(v // 2), v - v // 2, dtype=np.float64) for v in self.N]
        if self.c2r:
            ks[2] = np.arange(-(self.N[2] // 2), 1, dtype=np.float64)
        return ks

    def _k(self, t):
        shp = [1, 1, 1]
        shp[t] = -1
        return self._kvecs()[t].reshape(shp)

    def _fshape(self):
        return tuple(self.Nh) if self.c2r else tuple(self.N)

    def interlace_phase(self, sign):
        """convolution_due_to_interlacing (kernel/matrix_D.c:282-317): the factor
        e^{-sign pi i sum_t k_t/n_t}; trafo_D passes FFTW_FORWARD (-1), adjoint_D FFTW_BACKWARD (+1)."""
        ks = self._kvecs()
        h = ((ks[0] / self.n[0])[:, None, None] + (ks[1] / self.n[1])[None, :, None]
             + (ks[2] / self.n[2])[None, None, :])
        return np.exp(-sign * np.pi * 1j * h)

    def shifted_x(self, x):
        """Nodes of the interlaced pass: x_t + 1/(2 n_t), folded back into [-1/2, 1/2)
        (loop_over_particles_trafo / _adj, kernel/ndft-parallel.c:2732-2753, :2914-2934)."""
        xs = _f64(x).reshape(-1, 3) + 0.5 / np.asarray(self.n, np.float64)
        xs = np.where(xs >= 0.5, xs - 1.0, xs)
        return np.ascontiguousarray(xs.reshape(-1))

    def _D_any(self, fhat):
        if self.c2r:
            return self._D_half(fhat)
        return self.D(fhat).reshape(self.N)

    def _DT_any(self, g1, fhat):
        if self.c2r:
            fhat += self._D_half(g1).reshape(-1)
        else:
            self.DT_accumulate(np.asarray(g1).reshape(-1), fhat)

    def _grid_fwd(self, g1):
        """F and the ghost exchange of g1 (f_hat layout): the padded grid B reads."""
        if self.c2r:
            return np.ascontiguousarray(np.pad(self._F_c2r(g1), [(self.m, self.m)] * 3, mode="wrap"))
        return np.ascontiguousarray(self.ghost_exchange(self.F(np.asarray(g1).reshape(-1))))

    def _gather(self, x, gpad, f=None, grad=None, hess=None):
        """B at the nodes x (adds onto the given outputs)."""
        M = x.size // 3
        isr = 1 if self.c2r else 0

        def v(a):
            return None if a is None else _dp(a if isr else a.view(np.float64))

        fl = (PNFFT_COMPUTE_F if f is not None else 0) | (PNFFT_COMPUTE_GRAD_F if grad is not None else 0)
        args = (_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M, _dp(x),
                _dp(gpad if isr else gpad.view(np.float64)), isr)
        self._ip_on()
        if fl:
            lib().orc_B_trafo(*args, v(f), v(grad), fl)
        if hess is not None:
            lib().orc_B_trafo_hess(*args, v(hess))
        self._ip_off()

    def _grid_adj(self, x, f=None, grad=None):
        """B^T, ghost reduce and F^H of node values (f_hat layout)."""
        M = x.size // 3
        isr = 1 if self.c2r else 0
        ng = [v + 2 * self.m for v in self.no]
        gpad = np.zeros(ng, np.float64 if isr else np.complex128)
        fl = (PNFFT_COMPUTE_F if f is not None else 0) | (PNFFT_COMPUTE_GRAD_F if grad is not None else 0)

        def v(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64 if isr else np.complex128)
            return _dp(a if isr else a.view(np.float64)), a

        fv, gv = v(f), v(grad)
        self._ip_on()
        lib().orc_B_adj(_lp(_longs(self.n)), _lp(_longs(self.no)), self.m, self.window, _dp(self.b), M, _dp(x),
                        fv[0] if fv else None, gv[0] if gv else None, isr,
                        _dp(gpad if isr else gpad.view(np.float64)), fl)
        self._ip_off()
        g2 = self.ghost_reduce(gpad)
        return self._FH_r2c(g2) if self.c2r else self.FH(g2).reshape(self.N)

    def trafo_general(self, x, fhat, compute_f=True, compute_grad=False, compute_hess=False):
        """pnfft_trafo with the plan's PNFFT_INTERLACED / PNFFT_DIFF_IK flags and Hessians
        (api/api-basic.c:100-244): returns (f, grad, hess), complex for c2c plans, real for c2r.

        Interlacing (api/api-basic.c:233-240): the average of the plain pass and a pass at the
        nodes shifted by half a cell with D followed by the phase e^{+pi i sum k_t/n_t}
        (kernel/matrix_D.c:227-251); the 1/2 of kernel/assign.c:688 is applied to g1 here.
        ik-differentiation (:100-167): f from B F D f_hat; grad_t from B F (-2 pi i k_t D f_hat);
        Hessian (a, b) from B F (-4 pi^2 k_a k_b D f_hat), all with the f-gather.  Without
        DIFF_IK the window-derivative gathers give grad and Hessian (kernel/assign.c:763-1027)."""
        x = _f64(x).reshape(-1)
        M = x.size // 3
        dt = np.float64 if self.c2r else np.complex128
        f = np.zeros(M, dt) if compute_f else None
        grad = np.zeros(3 * M, dt) if compute_grad else None
        hess = np.zeros(6 * M, dt) if compute_hess else None
        passes = [(False, 0.5), (True, 0.5)] if self.flags & PNFFT_INTERLACED else [(False, 1.0)]
        for il, sc in passes:
            xx = self.shifted_x(x) if il else x
            g1 = self._D_any(fhat) * sc
            if il:
                g1 = g1 * self.interlace_phase(-1)
            if not self.flags & PNFFT_DIFF_IK:
                self._gather(xx, self._grid_fwd(g1), f, grad, hess)
                continue
            if compute_f:
                self._gather(xx, self._grid_fwd(g1), f=f)
            if compute_grad:  # trafo_scale_ik_diff_c2c, kernel/ndft-parallel.c:3034-3054
                for t in range(3):
                    tmp = np.zeros(M, dt)
                    self._gather(xx, self._grid_fwd(-2j * np.pi * self._k(t) * g1), f=tmp)
                    grad[t::3] += tmp
            if compute_hess:  # trafo_scale_ik_diff2_c2c, :3056-3087
                for q, (a, b) in enumerate(HESS_PAIRS):
                    tmp = np.zeros(M, dt)
                    self._gather(xx, self._grid_fwd(-4.0 * np.pi ** 2 * self._k(a) * self._k(b) * g1), f=tmp)
                    hess[q::6] += tmp
        return f, grad, hess

    def adj_general(self, x, f=None, grad=None, fhat_in=None):
        """pnfft_adj with the plan's PNFFT_INTERLACED / PNFFT_DIFF_IK flags (api/api-basic.c:247-378):
        f_hat (+)= sum over passes of D^T [phase] (1/2) g1, with g1 = F^H B^T (f, grad) (plain) or
        F^H B^T f + sum_t 2 pi i k_t F^H B^T grad_t (ik, adjoint_scale_ik_diff_c2c :3012-3032)."""
        x = _f64(x).reshape(-1)
        shape = self._fshape()
        fhat = (np.zeros(int(np.prod(shape)), np.complex128) if fhat_in is None
                else _c128(fhat_in).reshape(-1).copy())
        passes = [(False, 0.5), (True, 0.5)] if self.flags & PNFFT_INTERLACED else [(False, 1.0)]
        for il, sc in passes:
            xx = self.shifted_x(x) if il else x
            if not self.flags & PNFFT_DIFF_IK:
                g1 = self._grid_adj(xx, f=f, grad=grad)
            else:
                g1 = np.zeros(shape, np.complex128)
                if f is not None:
                    g1 = g1 + self._grid_adj(xx, f=f)
                if grad is not None:
                    gr = np.asarray(grad).reshape(-1, 3)
                    for t in range(3):
                        g1 = g1 + 2j * np.pi * self._k(t) * self._grid_adj(xx, f=gr[:, t].copy())
            g1 = g1 * sc
            if il:
                g1 = g1 * self.interlace_phase(+1)
            self._DT_any(g1, fhat)
        return fhat


def sort_nodes(n, m, x):
    """Reference node sort (kernel/ndft-parallel.c:2121-2159): returns (keys, perm)."""
    x = _f64(x).reshape(-1)
    M = x.size // 3
    out = np.zeros(2 * max(M, 1), np.int64)
    lib().orc_sort_nodes(_lp(_longs(n)), int(m), M, _dp(x), _lp(out))
    out = out[: 2 * M].reshape(M, 2)
    return out[:, 0].copy(), out[:, 1].copy()


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
