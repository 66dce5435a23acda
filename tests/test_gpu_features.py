"""Interlacing, ik-differentiation, Hessians and the compute-flag corners on the
GPU, through the C interface, against the FP64 oracle.

Tolerances: rel-l2 <= 1e-12 for FP64 plans, <= 1e-5 for FP32 plans (f, grad,
f_hat), 1e-4 for FP32 Hessians (the second derivatives of the windows reach
n^2 times their values, so the FP32 rounding of the stencil sums is amplified
by the cancellation between stencil points).
"""
import sys

import numpy as np
import pytest
import torch

import pnfft_amd as PA
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64

pytestmark = pytest.mark.gpu

WIN = {"kb": 0, "gauss": PA.PNFFT_WINDOW_GAUSSIAN, "bspline": PA.PNFFT_WINDOW_BSPLINE,
       "sincpow": PA.PNFFT_WINDOW_SINC_POWER}
IL, IK = PA.PNFFT_INTERLACED, PA.PNFFT_DIFF_IK
F, G, H = PA.PNFFT_COMPUTE_F, PA.PNFFT_COMPUTE_GRAD_F, PA.PNFFT_COMPUTE_HESSIAN_F


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(seed, N, M, c2r, oracle):
    rng = SplitMix64(seed)
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    # the domain's lower edge, grid points and a half-cell point near the upper edge; all three
    # exact in FP32 (0.5 - 1e-9 would round to 0.5, outside [-1/2, 1/2)), none within rounding
    # of a stencil edge (where the reference's Kaiser-Bessel derivative formulas cancel)
    x[:3] = [[-0.5, -0.5, -0.5], [0.5 - 2.0 ** -7, 0.25, -0.125], [0.0, 0.0, 0.0]]
    if c2r:
        fh = oracle.semirandom_hermitian(N, half=True) * 1e-3
        fv = rng.normal(M)
        gv = rng.normal(3 * M)
    else:
        fh = rng.complex_normal(int(np.prod(N)))
        fv = rng.complex_normal(M)
        gv = rng.complex_normal(3 * M)
    return x, fh, fv, gv


def _run(dp, x, fh, fv, gv, trafo_flags, want_adj=True):
    dp.set_x(x)
    dp.set_f_hat(fh)
    dp.alloc_hessian()
    dp.trafo(trafo_flags)
    out = {"f": dp.f.cpu().numpy().copy(), "g": dp.grad.cpu().numpy().reshape(-1).copy(),
           "h": dp.hess.cpu().numpy().reshape(-1).copy()}
    if want_adj:
        dp.set_f(fv)
        dp.set_grad(gv)
        dp.adj(F)
        out["af"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
        dp.adj(G)
        out["ag"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
        dp.adj(F | G)
        out["afg"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    assert err is None, err
    return out


CASES = [
    # (flags, window, m, c2r)
    (IL, "kb", 4, False),
    (IK, "kb", 4, False),
    (IL | IK, "gauss", 3, False),
    (IL, "bspline", 3, True),
    (IK, "kb", 4, True),
    (IL | IK, "sincpow", 3, False),
]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("flags,wname,m,c2r", CASES)
def test_interlaced_ik_match_oracle(gpu, oracle, precision, flags, wname, m, c2r):
    say(f"case features flags={flags:#x} {wname} m={m} c2r={c2r} {precision}")
    N, n, M = (16, 8, 32), (32, 16, 64), 400
    x, fh, fv, gv = _inputs(31, N, M, c2r, oracle)
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    dp = DevicePlan(N, n, m, WIN[wname] | flags, M, precision=precision, c2r=c2r)
    out = _run(dp, x, fh, fv, gv, F | G | H)
    dp.close()
    P = oracle.Plan(N, n, m, WIN[wname] | flags, c2r=c2r)
    f_o, g_o, h_o = P.trafo_general(x, fh, True, True, True)
    tol = 1e-12 if precision == "double" else 1e-5
    htol = 1e-12 if precision == "double" else 1e-4
    errs = dict(f=rel(out["f"], f_o), g=rel(out["g"], g_o), h=rel(out["h"], h_o),
                af=rel(out["af"], P.adj_general(x, f=fv)), ag=rel(out["ag"], P.adj_general(x, grad=gv)),
                afg=rel(out["afg"], P.adj_general(x, f=fv, grad=gv)))
    say("  rel-l2", errs)
    assert all(v < tol for k, v in errs.items() if k != "h"), errs
    assert errs["h"] < htol, errs


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("wname,m", [("kb", 4), ("gauss", 5), ("bspline", 3), ("sincpow", 4), ("kb", 7)])
def test_hessian_gather_matches_oracle(gpu, oracle, precision, wname, m):
    """PNFFT_COMPUTE_HESSIAN_F by window second derivatives (kernel/assign.c:881-1027),
    alone and with f and grad in the same call, plain and accumulated."""
    if precision == "double" and m >= 7:
        pytest.skip("FP64 plans are exercised at m <= 6 here; m = 7 runs in FP32")
    say(f"case hessian {wname} m={m} {precision}")
    N, n, M = (16, 8, 16), (32, 16, 32), 500
    x, fh, _, _ = _inputs(32, N, M, False, oracle)
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    dp = DevicePlan(N, n, m, WIN[wname], M, precision=precision)
    out = _run(dp, x, fh, None, None, H, want_adj=False)
    h1 = out["h"]
    dp.trafo(F | G | H | PA.PNFFT_COMPUTE_ACCUMULATED)
    h2 = dp.hess.cpu().numpy().reshape(-1)
    f2 = dp.f.cpu().numpy()
    assert dp.api.last_error() is None
    dp.close()
    P = oracle.Plan(N, n, m, WIN[wname])
    f_o, _, h_o = P.trafo_general(x, fh, True, False, True)
    tol = 1e-12 if precision == "double" else 1e-4
    e1, e2, ef = rel(h1, h_o), rel(h2, 2 * h_o), rel(f2, f_o)
    say(f"  rel-l2 hessian {e1:.2e}, accumulated {e2:.2e}, f {ef:.2e}")
    assert e1 < tol and e2 < tol and ef < (1e-12 if precision == "double" else 1e-5)


@pytest.mark.parametrize("c2r", [False, True])
def test_direct_hessian_matches_oracle_ndft(gpu, oracle, c2r):
    N, M = (8, 4, 16), 300
    x, fh, _, _ = _inputs(33, N, M, c2r, oracle)
    dp = DevicePlan(N, None, 3, 0, M, c2r=c2r)
    out = _run(dp, x, fh, None, None, F | G | H | PA.PNFFT_COMPUTE_DIRECT, want_adj=False)
    dp.close()
    ref = oracle.ndft_hessian_c2r(N, x, fh) if c2r else oracle.ndft_hessian(N, x, fh)
    assert rel(out["h"], ref) < 1e-12


def test_real_f_writes_real_parts(gpu, oracle):
    """PNFFT_REAL_F: the gathers read the grid's real part and write only the real parts
    of f, grad_f and hessian_f (assign_*_r2r, kernel/assign.c:715-739, :821-851)."""
    N, n, M, m = (8, 8, 8), (16, 16, 16), 200, 3
    x, fh, _, _ = _inputs(34, N, M, False, oracle)
    dp = DevicePlan(N, n, m, PA.PNFFT_REAL_F, M)
    dp.set_x(x)
    dp.set_f_hat(fh)
    dp.alloc_hessian()
    dp.f.fill_(complex(0, 5))
    dp.grad.fill_(complex(0, 7))
    dp.hess.fill_(complex(0, 9))
    dp.trafo(F | G | H | PA.PNFFT_COMPUTE_ACCUMULATED)
    f = dp.f.cpu().numpy()
    g = dp.grad.cpu().numpy().reshape(-1)
    h = dp.hess.cpu().numpy().reshape(-1)
    assert dp.api.last_error() is None
    dp.close()
    P = oracle.Plan(N, n, m, 0)
    f_o, g_o, h_o = P.trafo_general(x, fh, True, True, True)
    # Re(B G) = B Re(G) for real windows: the real parts equal those of the complex transform
    assert rel(f.real, f_o.real) < 1e-12 and np.all(f.imag == 5)
    assert rel(g.real, g_o.real) < 1e-12 and np.all(g.imag == 7)
    assert rel(h.real, h_o.real) < 1e-12 and np.all(h.imag == 9)


def test_omit_flags(gpu, oracle):
    """PNFFT_OMIT_FFT gathers from the grid the previous call left; PNFFT_OMIT_CONV stops
    after F; PNFFT_OMIT_DECONV transforms f_hat without D (api/api-basic.c:170-196)."""
    N, n, M, m = (8, 8, 8), (16, 16, 16), 150, 4
    x, fh, _, _ = _inputs(35, N, M, False, oracle)
    P = oracle.Plan(N, n, m, 0)
    f_o, _ = P.trafo(x, fh)
    dp = DevicePlan(N, n, m, 0, M)
    dp.set_x(x)
    dp.set_f_hat(fh)
    dp.trafo(F | PA.PNFFT_OMIT_CONV)          # grid = F D f_hat, f untouched (zeroed only with conv)
    dp.set_f_hat(np.zeros_like(fh))           # OMIT_FFT must not read f_hat
    dp.trafo(F | PA.PNFFT_OMIT_FFT)           # B of the grid left by the first call
    f1 = dp.f.cpu().numpy().copy()
    assert rel(f1, f_o) < 1e-12
    # OMIT_DECONV: B F f_hat, the reference's transform of g1 = f_hat
    dp.set_f_hat(fh)
    dp.trafo(F | PA.PNFFT_OMIT_DECONV)
    f2 = dp.f.cpu().numpy().copy()
    assert dp.api.last_error() is None
    dp.close()
    gpad = P.ghost_exchange(P.F(np.asarray(fh)))
    f_ref = np.zeros(M, np.complex128)
    P._gather(np.ascontiguousarray(x.reshape(-1)), np.ascontiguousarray(gpad), f=f_ref)
    assert rel(f2, f_ref) < 1e-12


def test_set_b_changes_the_window_shape(gpu, oracle):
    """set_b (api/api-basic.c:587-596) replaces the Kaiser-Bessel shape parameter per axis: the
    plan's D factors, window scale and windows follow it, so trafo (f, grad) and adj match the
    oracle run with the same b (its D tables rebuilt from inv_phi_hat with that b), get_b
    returns it, and the outputs differ from those of the default b."""
    import ctypes
    N, n, M, m = (8, 16, 8), (16, 32, 16), 200, 4
    x, fh, fv, _ = _inputs(41, N, M, False, oracle)
    P = oracle.Plan(N, n, m, 0)
    f_def, _ = P.trafo(x, fh)
    b_new = np.array([0.75 * P.b[0], 1.1 * P.b[1], 0.7 * P.b[2]])
    P.b = b_new
    P.tab = np.concatenate([[oracle.inv_phi_hat(P.window, i - N[t] // 2, n[t], b_new[t], m) for i in range(N[t])]
                            for t in range(3)])
    f_o, g_o = P.trafo(x, fh, compute_grad=True)
    a_o = P.adj(x, f=fv)
    assert rel(f_o, f_def) > 1e-6  # the new shape changes the transform beyond the tolerance below
    say("case set_b", b_new)
    dp = DevicePlan(N, n, m, 0, M)
    dp.api.set_b(float(b_new[0]), float(b_new[1]), float(b_new[2]), dp.plan)
    got = [ctypes.c_double() for _ in range(3)]
    dp.api.get_b(dp.plan, *[ctypes.byref(v) for v in got])
    assert [v.value for v in got] == list(b_new)
    out = _run(dp, x, fh, fv, None, F | G, want_adj=False)
    f, g = out["f"], out["g"]
    dp.set_f(fv)
    dp.adj(F)
    a = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    assert rel(f, f_o) < 1e-12
    assert rel(g, g_o) < 1e-12
    assert rel(a, a_o) < 1e-12


def test_interlaced_vjp_is_rejected(gpu):
    dp = DevicePlan((8, 8, 8), None, 3, IL, 16)
    w = torch.zeros(16, dtype=torch.complex128, device="cuda:0")
    with pytest.raises(RuntimeError):
        dp.grad_x_vjp(w)
        dp._check()
    dp.close()
