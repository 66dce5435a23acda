"""Writes the committed fixtures tests/golden/nfft_small.npz.

Small seeded cases (every window, two geometries): inputs, the direct NDFT
sums computed here with numpy from the definitions (api/pnfft.h:214-236 and
kernel/ndft-parallel.c:377-722: f_j = sum_k fhat_k e^{-2 pi i k.x_j}, the
gradient -2 pi i k times the same terms, the adjoint with e^{+2 pi i k.x_j}),
and the FP64 oracle's fast transforms of the same inputs.  The NDFT arrays
depend only on the mathematics; the fast-transform arrays pin the oracle
against later changes, and the GPU tests compare the HIP path with both
without rebuilding anything.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pnfft_amd.workloads import SplitMix64  # noqa: E402

WINDOWS = {"kb": 0, "gauss": 1 << 13, "bspline": 1 << 14, "sincpow": 1 << 15}
CASES = [  # name, N, n, m, window
    ("kb_m3", (8, 8, 8), (16, 16, 16), 3, "kb"),
    ("gauss_m3", (8, 8, 8), (16, 16, 16), 3, "gauss"),
    ("bspline_m3", (8, 8, 8), (16, 16, 16), 3, "bspline"),
    ("sincpow_m3", (8, 8, 8), (16, 16, 16), 3, "sincpow"),
    ("kb_m6_aniso", (8, 16, 4), (16, 32, 16), 6, "kb"),
]
M = 48


def freqs(N):
    return [np.arange(-(t // 2), t - t // 2) for t in N]


def ndft(N, x, fhat, f, g):
    """Direct sums: trafo f and grad, adjoint of f and of grad (row-major k, k2 fastest)."""
    k0, k1, k2 = np.meshgrid(*freqs(N), indexing="ij")
    K = np.stack([k0.ravel(), k1.ravel(), k2.ravel()], axis=1).astype(np.float64)  # (Ntot, 3)
    E = np.exp(-2j * np.pi * (x @ K.T))  # (M, Ntot)
    tf = E @ fhat
    tg = np.stack([E @ (fhat * (-2j * np.pi * K[:, t])) for t in range(3)], axis=1).reshape(-1)
    af = E.conj().T @ f
    gg = g.reshape(-1, 3)
    ag = sum((E.conj().T @ gg[:, t]) * (2j * np.pi * K[:, t]) for t in range(3))
    return tf, tg, af, ag


def main():
    out = {}
    for ci, (name, N, n, m, wn) in enumerate(CASES):
        rng = SplitMix64(777000 + ci)
        x = rng.uniform(3 * M).reshape(M, 3) - 0.5
        fhat = rng.complex_normal(int(np.prod(N)))
        f = rng.complex_normal(M)
        g = rng.complex_normal(3 * M)
        tf, tg, af, ag = ndft(N, x, fhat, f, g)
        # the oracle's own direct sums agree with the definitions above
        of, og = O.ndft_trafo(N, x, fhat, compute_grad=True)
        assert np.allclose(of, tf, rtol=0, atol=1e-10 * np.abs(tf).max())
        assert np.allclose(og, tg, rtol=0, atol=1e-10 * np.abs(tg).max())
        P = O.Plan(N, n, m, WINDOWS[wn])
        nf, ng = P.trafo(x, fhat, compute_grad=True)
        naf = P.adj(x, f=f)
        nag = P.adj(x, grad=g)
        rec = dict(N=np.array(N), n=np.array(n), m=np.array(m), flags=np.array(WINDOWS[wn]), x=x, fhat=fhat, f=f,
                   g=g, ndft_f=tf, ndft_grad=tg, ndft_adj_f=af, ndft_adj_grad=ag, nfft_f=nf, nfft_grad=ng,
                   nfft_adj_f=naf, nfft_adj_grad=nag)
        for k, v in rec.items():
            out[f"{name}__{k}"] = v
        err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in ((nf, tf), (ng, tg), (naf, af), (nag, ag)))
        print(f"{name}: fast vs direct rel-l2 {err:.2e}")
    np.savez_compressed(os.path.join(HERE, "nfft_small.npz"), **out)
    print("cases:", ", ".join(c[0] for c in CASES))


if __name__ == "__main__":
    main()
