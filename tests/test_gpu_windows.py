"""The remaining windows and window modes on the GPU, through the C interface,
against the FP64 oracle: the Bessel-I0 window, the truncated Gaussian
(PNFFT_WINDOW_GAUSSIAN_T), the interpolated windows of
PNFFT_PRE_{CONST,LIN,QUAD,CUB}_PSI and PNFFT_FAST_GAUSSIAN.

Tolerances: rel-l2 <= 1e-12 for FP64 plans, <= 1e-5 for FP32 plans (f, grad,
f_hat), 1e-4 for FP32 Hessians.
"""
import sys

import numpy as np
import pytest
import torch

import pnfft_amd as PA
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64

pytestmark = pytest.mark.gpu

F, G, H = PA.PNFFT_COMPUTE_F, PA.PNFFT_COMPUTE_GRAD_F, PA.PNFFT_COMPUTE_HESSIAN_F
BESSEL = PA.PNFFT_WINDOW_BESSEL_I0
GAUSS = PA.PNFFT_WINDOW_GAUSSIAN
GAUSS_T = PA.PNFFT_WINDOW_GAUSSIAN | (1 << 17)
PRE = {"const": 1 << 2, "lin": 1 << 3, "quad": 1 << 4, "cub": 1 << 5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _compare(oracle, flags, m, precision, c2r=False, seed=51, N=(16, 8, 16), n=(32, 16, 32), M=400,
             hess=True, real_parts=False):
    rng = SplitMix64(seed)
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    x[:2] = [[-0.5, 0.25, 0.0], [0.5 - 2.0 ** -7, -0.125, 0.375]]
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    if c2r:
        fh = oracle.semirandom_hermitian(N, half=True) * 1e-3
        fv, gv = rng.normal(M), rng.normal(3 * M)
    else:
        fh = rng.complex_normal(int(np.prod(N)))
        fv, gv = rng.complex_normal(M), rng.complex_normal(3 * M)
    dp = DevicePlan(N, n, m, flags, M, precision=precision, c2r=c2r)
    dp.set_x(x)
    dp.set_f_hat(fh)
    dp.alloc_hessian()
    dp.trafo(F | G | (H if hess else 0))
    f = dp.f.cpu().numpy().copy()
    g = dp.grad.cpu().numpy().reshape(-1).copy()
    h = dp.hess.cpu().numpy().reshape(-1).copy()
    dp.set_f(fv)
    dp.set_grad(gv)
    dp.adj(F | G)
    a = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    P = oracle.Plan(N, n, m, flags, c2r=c2r)
    f_o, g_o, h_o = P.trafo_general(x, fh, True, True, hess)
    a_o = P.adj_general(x, f=fv, grad=gv)
    if real_parts:  # PNFFT_PRE_QUAD_PSI has the value of PNFFT_REAL_F: real parts only (quirk)
        assert np.all(f.imag == 0) and np.all(g.imag == 0)
        f, g, h = f.real, g.real, h.real
        f_o, g_o = f_o.real, g_o.real
        h_o = h_o.real if hess else None
    tol = 1e-12 if precision == "double" else 1e-5
    htol = 1e-12 if precision == "double" else 1e-4
    errs = dict(f=rel(f, f_o), g=rel(g, g_o), a=rel(a, a_o))
    if hess:
        errs["h"] = rel(h, h_o)
    say("  rel-l2", errs)
    assert errs["f"] < tol and errs["g"] < tol and errs["a"] < tol, errs
    if hess:
        assert errs["h"] < htol, errs


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("m,extra,c2r", [(4, 0, False), (6, 0, False), (3, PA.PNFFT_INTERLACED, False),
                                         (4, PA.PNFFT_DIFF_IK, True)])
def test_bessel_i0_window(gpu, oracle, precision, m, extra, c2r):
    say(f"case bessel-i0 m={m} flags={extra:#x} c2r={c2r} {precision}")
    _compare(oracle, BESSEL | extra, m, precision, c2r=c2r)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("m", [3, 6])
def test_truncated_gaussian(gpu, oracle, precision, m):
    say(f"case gaussian_t m={m} {precision}")
    _compare(oracle, GAUSS_T, m, precision)


def test_truncated_gaussian_phi_hat_api(gpu, oracle):
    N, n, m = (16, 16, 16), (32, 32, 32), 5
    dp = DevicePlan(N, n, m, GAUSS_T, 8)
    tab = oracle.inv_phi_hat_tables(N, n, m, GAUSS_T)
    for t in range(3):
        for i in (0, 5, 8, 15):
            k = i - 8
            got = dp.api.inv_phi_hat(dp.plan, t, k)
            assert abs(got - tab[16 * t + i]) <= 1e-13 * abs(tab[16 * t + i])
            assert abs(dp.api.phi_hat(dp.plan, t, k) * tab[16 * t + i] - 1) < 1e-13
    dp.close()


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("key,wflags,m", [("const", 0, 4), ("lin", GAUSS, 3), ("quad", 0, 5), ("cub", 0, 4),
                                          ("cub", BESSEL, 3), ("cub", PA.PNFFT_DIFF_IK, 4)])
def test_interpolated_windows(gpu, oracle, precision, key, wflags, m):
    say(f"case intpol {key} flags={wflags:#x} m={m} {precision}")
    _compare(oracle, PRE[key] | wflags, m, precision, real_parts=(key == "quad"))


def test_fast_gaussian_equals_gaussian(gpu, oracle):
    """PNFFT_FAST_GAUSSIAN evaluates the same window by a recurrence (kernel/ndft-parallel.c:1694-1714);
    the exact evaluation here matches the oracle's plain Gaussian."""
    _compare(oracle, GAUSS | PA.PNFFT_FAST_GAUSSIAN, 4, "double")


def test_vjp_rejects_table_windows(gpu):
    dp = DevicePlan((8, 8, 8), None, 3, BESSEL, 16)
    w = torch.zeros(16, dtype=torch.complex128, device="cuda:0")
    dp.grad_x_vjp(w)
    assert dp.api.last_error() is not None
    dp.close()
