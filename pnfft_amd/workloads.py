"""Seeded synthetic NFFT workloads (nodes x, coefficients f_hat, samples f).

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014): state advances
by 0x9E3779B97F4A7C15 and each output is the state passed through the
finalizer z ^= z>>30; z *= 0xBF58476D1CE4E5B9; z ^= z>>27;
z *= 0x94D049BB133111EB; z ^= z>>31.  Uniform doubles are (z >> 11) * 2^-53;
normals use Box-Muller on pairs (u1, u2): sqrt(-2 ln(1-u1)) * (cos, sin)(2 pi u2).

The five workloads are the node sets and sizes the NFFT is measured and
checked on; each takes a ``scale`` for M so that tests can run a cut-down
node count with the same distribution.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * GOLDEN
            self.state = self.state + np.uint64(n) * GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _C1
            z = (z ^ (z >> np.uint64(27))) * _C2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n):
        """Doubles in [0, 1)."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        k = (n + 1) // 2
        u = self.uniform(2 * k).reshape(k, 2)
        r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
        th = 2.0 * np.pi * u[:, 1]
        return np.stack([r * np.cos(th), r * np.sin(th)], axis=1).reshape(-1)[:n]

    def complex_normal(self, n):
        v = self.normal(2 * n).reshape(n, 2)
        return v[:, 0] + 1j * v[:, 1]

    def permutation(self, n):
        keys = self.next_u64(n)
        return np.argsort(keys, kind="stable")


def wrap(x):
    """Map onto the torus [-1/2, 1/2)."""
    y = x - np.floor(x + 0.5)
    y[y >= 0.5] -= 1.0
    return y


# window flags (include/pnfft.h)
_KB, _GAUSS, _BSPLINE = 0, 1 << 13, 1 << 14

CONFIGS = {
    1: dict(name="kb_m6_N16_M4096", N=(16, 16, 16), n=(32, 32, 32), m=6, flags=_KB, M=4096, nodes="uniform",
            precision="double"),
    2: dict(name="kb_m6_N128_M2^22", N=(128, 128, 128), n=(256, 256, 256), m=6, flags=_KB, M=1 << 22,
            nodes="uniform", precision="double"),
    3: dict(name="kb_m4_N256_M2^24_gaussmix", N=(256, 256, 256), n=(512, 512, 512), m=4, flags=_KB, M=1 << 24,
            nodes="gaussmix", precision="float"),
    4: dict(name="gauss_m8_N64_M2^20_c2r", N=(64, 64, 64), n=(128, 128, 128), m=8, flags=_GAUSS, M=1 << 20,
            nodes="uniform", precision="double", c2r=True),
    5: dict(name="bspline_m6_N256x256x64_M2^25_slab", N=(256, 256, 64), n=(512, 512, 128), m=6, flags=_BSPLINE,
            M=1 << 25, nodes="slab", precision="float"),
}


def nodes(kind, M, rng):
    if kind == "uniform":
        return rng.uniform(3 * M).reshape(M, 3) - 0.5
    if kind == "gaussmix":
        centers = rng.uniform(24).reshape(8, 3) - 0.5
        which = np.arange(M) % 8
        which = which[rng.permutation(M)]
        x = centers[which] + 0.05 * rng.normal(3 * M).reshape(M, 3)
        return wrap(x)
    if kind == "slab":
        h = M // 2
        a = rng.uniform(3 * h).reshape(h, 3) - 0.5
        b = rng.uniform(3 * (M - h)).reshape(M - h, 3) - 0.5
        b[:, 2] = 0.04 * (b[:, 2] + 0.5) - 0.02
        return np.concatenate([a, b], axis=0)
    raise ValueError(kind)


def make(cfg, scale=1.0, M=None, seed=None):
    """Inputs of workload `cfg` (1..5): dict with N, n, m, flags, x (M,3) float64,
    f_hat (prod N,) complex128, f (M,) complex128."""
    c = dict(CONFIGS[cfg])
    if M is None:
        M = max(1, int(round(c["M"] * scale)))
    rng = SplitMix64(20261200 + cfg if seed is None else seed)
    x = nodes(c["nodes"], M, rng)
    Ntot = int(np.prod(c["N"]))
    fh = rng.complex_normal(Ntot)
    f = rng.complex_normal(M)
    c.update(M=M, x=np.ascontiguousarray(x), f_hat=fh, f=f)
    return c
