"""Autograd harness of the NFFT (GPU): backward products against central finite
differences and against the reference's PNFFT_COMPUTE_GRAD_F output."""
import sys

import numpy as np
import pytest
import torch

import pnfft_amd as PA
from pnfft_amd.autograd import nfft
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64

pytestmark = pytest.mark.gpu


def _setup(precision="double", M=64, N=(8, 8, 8), n=(16, 16, 16), m=4, flags=0, seed=31):
    rng = SplitMix64(seed)
    dev = "cuda:0"
    dt = torch.float64 if precision == "double" else torch.float32
    ct = torch.complex128 if precision == "double" else torch.complex64
    x = torch.tensor(rng.uniform(3 * M).reshape(M, 3) - 0.5, dtype=dt, device=dev)
    fh = torch.tensor(rng.complex_normal(int(np.prod(N))).reshape(N), dtype=ct, device=dev)
    w = torch.tensor(rng.complex_normal(M), dtype=ct, device=dev)
    plan = DevicePlan(N, n, m, flags, M, precision=precision)
    return plan, x, fh, w


def _loss(plan, x, fh, w):
    return torch.real(torch.sum(torch.conj(w) * nfft(x, fh, plan)))


@pytest.mark.parametrize("flags", [0, PA.PNFFT_WINDOW_GAUSSIAN])
def test_backward_matches_finite_differences(gpu, flags):
    print("case autograd fd", flags, flush=True)
    plan, x, fh, w = _setup(flags=flags)
    xr = x.clone().requires_grad_(True)
    fr = fh.clone().requires_grad_(True)
    L = _loss(plan, xr, fr, w)
    L.backward()
    gx, gf = xr.grad.clone(), fr.grad.clone()
    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(6):
        j, t = int(rng.integers(x.shape[0])), int(rng.integers(3))
        xp, xm = x.clone(), x.clone()
        xp[j, t] += h
        xm[j, t] -= h
        fd = (float(_loss(plan, xp, fh, w)) - float(_loss(plan, xm, fh, w))) / (2 * h)
        assert abs(fd - float(gx[j, t])) <= 1e-6 * max(1.0, abs(fd)), (j, t, fd, float(gx[j, t]))
    flat = fh.reshape(-1)
    for _ in range(6):
        k = int(rng.integers(flat.numel()))
        for unit in (1.0, 1j):
            fp, fm = flat.clone(), flat.clone()
            fp[k] += h * unit
            fm[k] -= h * unit
            fd = (float(_loss(plan, x, fp.reshape(fh.shape), w)) - float(_loss(plan, x, fm.reshape(fh.shape), w))) / (2 * h)
            got = float(gf.reshape(-1)[k].real if unit == 1.0 else gf.reshape(-1)[k].imag)
            assert abs(fd - got) <= 1e-6 * max(1.0, abs(fd)), (k, unit, fd, got)
    plan.close()


def test_grad_x_equals_compute_grad_f_contraction(gpu):
    """d/dx of Re<w, f> is Re(conj(w) grad_f) with grad_f from pnfft_trafo(PNFFT_COMPUTE_GRAD_F)."""
    print("case autograd vs grad_f", flush=True)
    plan, x, fh, w = _setup(M=500, N=(16, 16, 16), n=(32, 32, 32), m=6)
    xr = x.clone().requires_grad_(True)
    _loss(plan, xr, fh, w).backward()
    plan.set_x(x)
    plan.set_f_hat(fh)
    plan.trafo(PA.PNFFT_COMPUTE_GRAD_F)
    ref = torch.real(torch.conj(w)[:, None] * plan.grad)
    err = float(torch.linalg.norm(xr.grad - ref) / torch.linalg.norm(ref))
    assert err < 1e-13, err
    plan.close()


def test_fp32_backward_close_to_fp64(gpu):
    print("case autograd fp32", flush=True)
    outs = []
    for prec in ("double", "float"):
        plan, x, fh, w = _setup(precision=prec, M=300, N=(16, 16, 16), n=(32, 32, 32), m=4, seed=5)
        xr = x.clone().requires_grad_(True)
        fr = fh.clone().requires_grad_(True)
        _loss(plan, xr, fr, w).backward()
        outs.append((xr.grad.double().cpu(), fr.grad.to(torch.complex128).cpu()))
        plan.close()
    for a, b in zip(outs[1], outs[0]):
        assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) < 1e-5
