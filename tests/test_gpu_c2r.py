"""c2r plans (pnfft_init_guru_c2r) on the GPU against the FP64 oracle.

The half spectrum holds k2 in [-N2/2, 0]; f and grad_f are real.  Inputs with
zero Nyquist planes (the reference's own semirandom test input, and seeded
1/|k|^2-weighted fields) keep the NDFT and the padded c2r transform equal.
"""
import numpy as np
import pytest

import pnfft_amd as PA
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64, make

pytestmark = pytest.mark.gpu

# An FP64 Gaussian m = 8 c2r run stalled the GPU under an earlier kernel build
# (static-width FP64 m = 7, 8 kernels, since reverted).  No cause was shown, so
# the library refuses FP64 c2r plans with m >= 7 and these cases stay skipped.
FP64_M8_STALL = ("FP64 c2r m=8 stalled a GPU run of an earlier kernel build with no cause shown; "
                 "such plans are refused (pnfft_init_guru_c2r returns NULL)")


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run(N, n, m, flags, x, fh_half, fv, gv=None, precision="double", direct=False):
    M = len(x)
    dp = DevicePlan(N, n, m, flags, M, precision=precision, c2r=True)
    dp.set_x(x)
    dp.set_f_hat(fh_half)
    extra = PA.PNFFT_COMPUTE_DIRECT if direct else 0
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F | extra)
    out = {"f": dp.f.cpu().numpy().copy(), "grad": dp.grad.cpu().numpy().reshape(-1).copy()}
    dp.set_f(fv)
    dp.adj(PA.PNFFT_COMPUTE_F | extra)
    out["adj"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
    if gv is not None:
        dp.set_grad(gv)
        dp.adj(PA.PNFFT_COMPUTE_GRAD_F | extra)
        out["adj_grad"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    return out


def coulomb_half(N, rng):
    """Real field: N(0,1) half-spectrum coefficients times 1/|k|^2, k = 0 and Nyquist planes zero."""
    h = (N[0], N[1], N[2] // 2 + 1)
    c = rng.complex_normal(int(np.prod(h))).reshape(h)
    k0 = np.arange(-(N[0] // 2), N[0] - N[0] // 2)[:, None, None]
    k1 = np.arange(-(N[1] // 2), N[1] - N[1] // 2)[None, :, None]
    k2 = np.arange(-(N[2] // 2), 1)[None, None, :]
    k2sum = k0 ** 2 + k1 ** 2 + k2 ** 2
    with np.errstate(divide="ignore"):
        w = np.where(k2sum > 0, 1.0 / np.maximum(k2sum, 1), 0.0)
    nyq = (k0 == -(N[0] // 2)) | (k1 == -(N[1] // 2)) | (k2 == -(N[2] // 2))
    c = c * w
    c[np.broadcast_to(nyq, h)] = 0
    return c


@pytest.mark.parametrize("wname,m", [("kb", 6), pytest.param("gauss", 8, marks=pytest.mark.skip(reason=FP64_M8_STALL))])
def test_c2r_fp64_matches_oracle(gpu, oracle, wname, m):
    print("case c2r fp64", wname, flush=True)
    flags = {"kb": 0, "gauss": PA.PNFFT_WINDOW_GAUSSIAN}[wname]
    N, n = (16, 16, 16), (32, 32, 32)
    rng = SplitMix64(41)
    M = 1500
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    fh = oracle.semirandom_hermitian(list(N), half=True)
    fv = rng.normal(M)
    gv = rng.normal(3 * M)
    g = _run(N, n, m, flags, x, fh, fv, gv)
    P = oracle.Plan(N, n, m, flags, c2r=True)
    f_o, g_o = P.trafo_c2r(x, fh, compute_grad=True)
    assert rel(g["f"], f_o) < 1e-12
    assert rel(g["grad"], g_o) < 1e-12
    assert rel(g["adj"], P.adj_c2r(x, f=fv)) < 1e-12
    assert rel(g["adj_grad"], P.adj_c2r(x, grad=gv)) < 1e-12


def test_c2r_direct_matches_reference_ndft(gpu, oracle):
    """PNFFT_COMPUTE_DIRECT on a c2r plan, including the reference NDFT's gradient sign."""
    print("case c2r direct", flush=True)
    N, n = (8, 8, 8), (16, 16, 16)
    rng = SplitMix64(42)
    M = 200
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    fh = oracle.semirandom_hermitian(list(N), half=True)
    fv = rng.normal(M)
    gv = rng.normal(3 * M)
    g = _run(N, n, 4, 0, x, fh, fv, gv, direct=True)
    f_nd, g_nd = oracle.ndft_trafo_c2r(list(N), x, fh, compute_grad=True)
    assert rel(g["f"], f_nd) < 1e-12
    assert rel(g["grad"], g_nd) < 1e-12
    assert rel(g["adj"], oracle.ndft_adj_c2r(list(N), x, f=fv)) < 1e-12
    assert rel(g["adj_grad"], oracle.ndft_adj_c2r(list(N), x, grad=gv)) < 1e-12


@pytest.mark.parametrize("m", [6, 4, 3])
def test_c2r_fp32(gpu, oracle, m):
    print("case c2r fp32 m", m, flush=True)
    N, n = (16, 16, 16), (32, 32, 32)
    rng = SplitMix64(43)
    M = 1000
    x = (rng.uniform(3 * M).reshape(M, 3) - 0.5).astype(np.float32).astype(np.float64)
    fh = coulomb_half(N, rng)
    fv = rng.normal(M)
    g = _run(N, n, m, 0, x, fh, fv, precision="float")
    P = oracle.Plan(N, n, m, 0, c2r=True)
    f_o, g_o = P.trafo_c2r(x, fh, compute_grad=True)
    assert rel(g["f"], f_o) < 1e-5
    assert rel(g["grad"], g_o) < 1e-5
    assert rel(g["adj"], P.adj_c2r(x, f=fv)) < 1e-5


def test_c2r_particle_mesh_workload_fp32(gpu, oracle):
    """The particle-mesh geometry in FP32 (N=64^3, n=128^3, Gaussian m=8 on 4-cell tiles,
    charges +-1, 1/|k|^2 field): trafo (potential and field) and adj (charge spreading)."""
    print("case c2r particle mesh fp32", flush=True)
    w = make(4, M=1 << 14)
    rng = SplitMix64(20261214)
    N, n, m = w["N"], w["n"], w["m"]
    M = w["M"]
    x = w["x"].astype(np.float32).astype(np.float64)
    q = np.where(np.arange(M) % 2 == 0, 1.0, -1.0)[rng.permutation(M)]
    fh = coulomb_half(N, rng)
    g = _run(N, n, m, w["flags"], x, fh, q, precision="float")
    P = oracle.Plan(N, n, m, w["flags"], c2r=True)
    f_o, g_o = P.trafo_c2r(x, fh, compute_grad=True)
    assert rel(g["f"], f_o) < 1e-5
    assert rel(g["grad"], g_o) < 1e-5
    assert rel(g["adj"], P.adj_c2r(x, f=q)) < 1e-5


@pytest.mark.skip(reason=FP64_M8_STALL)
def test_c2r_particle_mesh_workload(gpu, oracle):
    """The particle-mesh geometry: N=64^3, n=128^3, Gaussian m=8, charges +-1, 1/|k|^2 field."""
    print("case c2r particle mesh", flush=True)
    w = make(4, M=1 << 14)
    rng = SplitMix64(20261204)
    N, n, m = w["N"], w["n"], w["m"]
    M = w["M"]
    q = np.where(np.arange(M) % 2 == 0, 1.0, -1.0)[rng.permutation(M)]
    fh = coulomb_half(N, rng)
    g = _run(N, n, m, w["flags"], w["x"], fh, q)
    P = oracle.Plan(N, n, m, w["flags"], c2r=True)
    f_o, g_o = P.trafo_c2r(w["x"], fh, compute_grad=True)
    assert rel(g["f"], f_o) < 1e-12
    assert rel(g["grad"], g_o) < 1e-12
    assert rel(g["adj"], P.adj_c2r(w["x"], f=q)) < 1e-12


@pytest.mark.parametrize("precision,N,n,m", [("double", (8, 64, 256), (16, 128, 512), 4),
                                              ("float", (8, 64, 256), (16, 128, 512), 4),
                                              ("double", (8, 8, 16), (16, 16, 16), 3)])
def test_c2r_real_axis_transforms(gpu, oracle, precision, N, n, m):
    """The real grid of a c2r plan (axis 2 as a complex transform of half length): a 512-point
    axis (256-point half transform) and an unpadded axis 2 (n2 = N2, whose k2 = -N2/2 plane is the
    Nyquist plane of the real transform), trafo and adjoint against the oracle.  The unpadded grid
    runs in FP64 only: without oversampling the deconvolution factors near the band edge are large
    enough to lift FP32 rounding above 1e-5 (9e-5 measured for the adjoint)."""
    rng = SplitMix64(93)
    M = 1200
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    fh = oracle.semirandom_hermitian(N, half=True) * 1e-3
    fv, gv = rng.normal(M), rng.normal(3 * M)
    out = _run(N, n, m, 0, x, fh, fv, gv, precision=precision)
    P = oracle.Plan(N, n, m, 0, c2r=True)
    f_o, g_o = P.trafo_c2r(x, fh, compute_grad=True)
    tol = 1e-12 if precision == "double" else 1e-5
    errs = dict(f=rel(out["f"], f_o), g=rel(out["grad"], g_o), a=rel(out["adj"], P.adj_c2r(x, f=fv)),
                ag=rel(out["adj_grad"], P.adj_c2r(x, grad=gv)))
    assert all(v < tol for v in errs.values()), errs


def test_c2r_grid_is_the_real_transform(gpu, oracle):
    """The oversampled grid of a c2r plan holds n0 n1 n2 reals: with PNFFT_OMIT_DECONV it equals
    Re(fftn) of the half spectrum with its k2 < 0 coefficients counted twice (numpy, FP64), and the
    adjoint's F^H of that real grid equals numpy's inverse transform at the stored frequencies."""
    import ctypes
    import torch
    N, n, m = (16, 8, 32), (32, 16, 64), 4
    H = N[2] // 2 + 1
    fh = oracle.semirandom_hermitian(N, half=True).reshape(N[0], N[1], H)
    dp = DevicePlan(N, n, m, 0, 0, precision="double", c2r=True)
    dp.set_f_hat(fh)
    dp.api.trafo(dp.plan, None, PA.PNFFT_OMIT_DECONV | PA.PNFFT_OMIT_CONV)
    assert dp.api.last_error() is None
    ptr = dp.api.oversampled_grid(dp.plan)
    grid = torch.empty(tuple(n), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(ptr), grid.numel() * 8, 3) == 0
    g = grid.cpu().numpy()
    e = sum(int(np.frexp(abs(dp.api.psi(dp.plan, t, 0.0)))[1]) - 1 for t in range(3))
    A = np.zeros(n, np.complex128)
    i0 = np.arange(-(N[0] // 2), N[0] // 2) % n[0]
    i1 = np.arange(-(N[1] // 2), N[1] // 2) % n[1]
    k2 = np.arange(-(N[2] // 2), 1)
    A[np.ix_(i0, i1, k2 % n[2])] = fh * np.where(k2 < 0, 2.0, 1.0)[None, None, :]
    ref = np.fft.fftn(A).real * 2.0 ** e
    assert rel(g, ref) < 1e-12
    # backward: f_hat = D^T F^H grid over the stored half, with the plan's own D factors
    dp.api.adj(dp.plan, None, PA.PNFFT_OMIT_CONV)
    assert dp.api.last_error() is None
    out = dp.f_hat.cpu().numpy().reshape(N[0], N[1], H)
    back = (np.fft.ifftn(g) * float(np.prod(n)))[np.ix_(i0, i1, k2 % n[2])]
    d = [np.array([dp.api.inv_phi_hat(dp.plan, t, k) for k in range(-(N[t] // 2), N[t] - N[t] // 2)])
         for t in range(3)]
    D = d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :H]
    dp.close()
    assert rel(out, back * D * 2.0 ** e) < 1e-12
