"""The measured workloads at their full sizes, HIP path against the FP64 oracle (GPU).

Each workload runs through the C interface exactly as the benchmark does
(device-resident arrays): trafo(f), trafo(grad f), adj(f).  The oracle
computes D and F once on the full grid and gathers f and grad f at a seeded
sample of the nodes (the gather is independent per node); the adjoint f_hat
depends on every node, so the oracle scatters all of them (OpenMP over the
host cores).  Tolerances: rel-l2 <= 1e-12 for the FP64 plans, <= 1e-5 for the
FP32 plans (against FP64, as the benchmark's accuracy gate).

The FFT tests compare the pruned pass chain with torch.fft (rocFFT) on the
device: trafo with PNFFT_OMIT_DECONV | PNFFT_OMIT_CONV leaves F f_hat (times the
plan's power-of-two window scale) in the plan's oversampled grid, and adj with
PNFFT_OMIT_CONV then gives D^T F^H of that grid.
"""
import ctypes
import sys
import time

import numpy as np
import pytest
import torch

import pnfft_amd as PA
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64, make

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def sample_nodes(M, k, seed):
    """k distinct node indices, seeded, sorted."""
    rng = SplitMix64(seed)
    idx = np.unique((rng.next_u64(2 * k) % np.uint64(M)).astype(np.int64))
    return idx[:k]


# workload -> (precompute flags used by the GPU run, tolerance)
CASES = {
    2: (0, 1e-12),                                        # plain method, windows evaluated in the kernels
    3: (PA.PNFFT_PRE_PSI | PA.PNFFT_PRE_GRAD_PSI, 1e-5),  # the benchmark's path: window tables
    5: (0, 1e-5),                                         # window recomputation (no tables)
}


@pytest.mark.parametrize("cfg", [2, 3, 5])
def test_workload_full_size_matches_oracle(gpu, oracle, cfg):
    pre, tol = CASES[cfg]
    t0 = time.time()
    w = make(cfg)
    prec = w["precision"]
    x = w["x"].astype(np.float32).astype(np.float64) if prec == "float" else w["x"]
    N, n, m, flags, M = w["N"], w["n"], w["m"], w["flags"], w["M"]
    say(f"case workload {w['name']} precision={prec} M={M} (inputs {time.time() - t0:.1f}s)")
    dp = DevicePlan(N, n, m, flags, M, precision=prec)
    dp.set_x(x)
    dp.set_f_hat(w["f_hat"])
    dp.set_f(w["f"])
    if pre:
        dp.precompute(pre)
    dp.trafo(PA.PNFFT_COMPUTE_F)
    f_gpu = dp.f.cpu().numpy().copy()
    dp.trafo(PA.PNFFT_COMPUTE_GRAD_F)
    g_gpu = dp.grad.cpu().numpy().copy()
    dp.set_f(w["f"])
    dp.adj(PA.PNFFT_COMPUTE_F)
    a_gpu = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    say(f"  gpu done ({time.time() - t0:.1f}s)")
    assert np.isfinite(f_gpu).all() and np.isfinite(g_gpu).all() and np.isfinite(a_gpu).all()

    P = oracle.Plan(N, n, m, flags)
    S = sample_nodes(M, 1 << 14, 20261230 + cfg)
    f_o, g_o = P.trafo(x[S], w["f_hat"], compute_grad=True)
    e_f = rel(f_gpu[S], f_o)
    e_g = rel(g_gpu[S].reshape(-1), g_o)
    say(f"  trafo sample of {len(S)}: rel-l2 f {e_f:.3e}, grad {e_g:.3e} ({time.time() - t0:.1f}s)")
    a_o = P.adj(x, f=w["f"], parallel=True)
    e_a = rel(a_gpu, a_o)
    say(f"  adj over all {M} nodes: rel-l2 {e_a:.3e} ({time.time() - t0:.1f}s)")
    assert e_f < tol and e_g < tol and e_a < tol, (e_f, e_g, e_a)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_long_axes_match_oracle(gpu, oracle, precision):
    """Grids with 512- and 256-point axes (the row pass of 512 points and the 256/512-point
    column passes), anisotropic, every window mode of the gather and the scatter."""
    N, n, m = (8, 128, 256), (16, 256, 512), 4
    say("case long axes", precision)
    rng = SplitMix64(77)
    M = 1500
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    fh = rng.complex_normal(int(np.prod(N)))
    fv = rng.complex_normal(M)
    gv = rng.complex_normal(3 * M)
    dp = DevicePlan(N, n, m, 0, M, precision=precision)
    dp.set_x(x)
    dp.set_f_hat(fh)
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
    f = dp.f.cpu().numpy().copy()
    g = dp.grad.cpu().numpy().reshape(-1).copy()
    dp.set_f(fv)
    dp.adj(PA.PNFFT_COMPUTE_F)
    a = dp.f_hat.cpu().numpy().reshape(-1).copy()
    dp.set_grad(gv)
    dp.adj(PA.PNFFT_COMPUTE_GRAD_F)
    ag = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    tol = 1e-12 if precision == "double" else 1e-5
    P = oracle.Plan(N, n, m, 0)
    f_o, g_o = P.trafo(x, fh, compute_grad=True)
    assert rel(f, f_o) < tol
    assert rel(g, g_o) < tol
    assert rel(a, P.adj(x, f=fv)) < tol
    assert rel(ag, P.adj(x, grad=gv)) < tol


def _copy_grid(dp, n, cdt):
    """The plan's oversampled grid (device) copied into a new torch tensor."""
    ptr = dp.api.oversampled_grid(dp.plan)
    assert ptr
    out = torch.empty(tuple(n), dtype=cdt, device="cuda:0")
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rc = hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), out.numel() * out.element_size(), 3)
    assert rc == 0
    return out


def _embed(fh, N, n):
    """f_hat (index k + N_t/2) placed at k mod n_t of a zero n-grid (torch, device)."""
    A = torch.zeros(tuple(n), dtype=fh.dtype, device=fh.device)
    idx = [torch.arange(-(Nt // 2), Nt - Nt // 2, device=fh.device) % nt for Nt, nt in zip(N, n)]
    A[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]] = fh
    return A, idx


@pytest.mark.parametrize("precision,N,n", [("double", (32, 64, 16), (64, 128, 32)),
                                           ("float", (64, 32, 128), (128, 64, 256)),
                                           ("float", (128, 64, 32), (512, 128, 64))])
def test_fft_chain_matches_rocfft(gpu, precision, N, n):
    say("case fft vs rocfft", precision, N, n)
    cdt = torch.complex128 if precision == "double" else torch.complex64
    rng = SplitMix64(91)
    fh_np = rng.complex_normal(int(np.prod(N))).reshape(N)
    dp = DevicePlan(N, n, 4, 0, 0, precision=precision)
    dp.set_f_hat(fh_np)
    dp.api.trafo(dp.plan, None, PA.PNFFT_OMIT_DECONV | PA.PNFFT_OMIT_CONV)
    assert dp.api.last_error() is None
    grid = _copy_grid(dp, n, cdt)
    fh = torch.tensor(fh_np, dtype=torch.complex128, device="cuda:0")
    A, idx = _embed(fh, N, n)
    # the plan's window values carry 2^-e per axis (e = exponent of psi(0)) and its D
    # factors 2^e; with PNFFT_OMIT_DECONV the transform keeps the 2^e, so that B F
    # is the reference's B F
    e = sum(int(np.frexp(abs(dp.api.psi(dp.plan, t, 0.0)))[1]) - 1 for t in range(3))
    ref = torch.fft.fftn(A) * 2.0 ** e  # rocFFT, FP64
    e_fwd = float(torch.linalg.norm(grid.to(torch.complex128) - ref) / torch.linalg.norm(ref))
    # backward: f_hat = D^T F^H grid with the plan's own D factors
    dp.api.adj(dp.plan, None, PA.PNFFT_OMIT_CONV)
    assert dp.api.last_error() is None
    out = dp.f_hat.to(torch.complex128)
    back = torch.fft.ifftn(grid.to(torch.complex128)) * float(np.prod(n))
    back = back[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
    d = [torch.tensor([dp.api.inv_phi_hat(dp.plan, t, k) for k in range(-(N[t] // 2), N[t] - N[t] // 2)],
                      dtype=torch.float64, device="cuda:0") for t in range(3)]
    D = d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :]
    # the D factors' 2^e stays in f_hat with PNFFT_OMIT_CONV
    ref_b = back * D * 2.0 ** e
    e_bwd = float(torch.linalg.norm(out - ref_b) / torch.linalg.norm(ref_b))
    dp.close()
    tol = 1e-12 if precision == "double" else 2e-6
    say(f"  rel-l2 forward {e_fwd:.3e}, backward {e_bwd:.3e}")
    assert e_fwd < tol and e_bwd < tol, (e_fwd, e_bwd)
