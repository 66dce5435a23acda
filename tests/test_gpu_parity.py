"""HIP path vs the FP64 oracle, through the C interface (GPU).

Tolerances: FP64 plans match the oracle's restatement of the reference
algorithm to rel-l2 <= 1e-12 (same algorithm, different FFT and summation
order); FP32 plans match it to rel-l2 <= 1e-5; both match the NDFT to the
window's accuracy.  Node binning is integer work and must equal numpy's stable
sort bit for bit.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

import pnfft_amd as PA
from pnfft_amd.device import DevicePlan
from pnfft_amd.workloads import SplitMix64, make

pytestmark = pytest.mark.gpu

WINDOWS = {"kb": 0, "gauss": 1 << 13, "bspline": 1 << 14, "sincpow": 1 << 15}
NDFT_ACC = {"kb": 1e-10, "gauss": 2e-5, "bspline": 8e-6, "sincpow": 3e-6}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_gpu(N, n, m, flags, x, fh, fv=None, gv=None, precision="double", grad=True, precompute=False):
    M = len(x)
    dp = DevicePlan(N, n, m, flags, M, precision=precision)
    dp.set_x(x)
    dp.set_f_hat(fh)
    if precompute:
        dp.precompute()
    out = {}
    dp.trafo(PA.PNFFT_COMPUTE_F | (PA.PNFFT_COMPUTE_GRAD_F if grad else 0))
    out["f"] = dp.f.cpu().numpy().copy()
    if grad:
        out["grad"] = dp.grad.cpu().numpy().reshape(-1).copy()
    if fv is not None:
        dp.set_f(fv)
        dp.adj(PA.PNFFT_COMPUTE_F)
        out["adj"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
    if gv is not None:
        dp.set_grad(gv)
        dp.adj(PA.PNFFT_COMPUTE_GRAD_F)
        out["adj_grad"] = dp.f_hat.cpu().numpy().reshape(-1).copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    return out


@pytest.mark.parametrize("wname", list(WINDOWS))
def test_fp64_matches_oracle_cfg1(gpu, oracle, wname):
    w = make(1)
    N, n, m = w["N"], w["n"], w["m"]
    flags = WINDOWS[wname]
    say("case fp64 cfg1", wname)
    rng = SplitMix64(7)
    gv = rng.complex_normal(3 * w["M"])
    g = run_gpu(N, n, m, flags, w["x"], w["f_hat"], w["f"], gv)
    P = oracle.Plan(N, n, m, flags)
    f_o, g_o = P.trafo(w["x"], w["f_hat"], compute_grad=True)
    assert rel(g["f"], f_o) < 1e-12
    assert rel(g["grad"], g_o) < 1e-12
    assert rel(g["adj"], P.adj(w["x"], f=w["f"])) < 1e-12
    assert rel(g["adj_grad"], P.adj(w["x"], grad=gv)) < 1e-12
    f_nd, _ = oracle.ndft_trafo(N, w["x"], w["f_hat"])
    assert rel(g["f"], f_nd) < NDFT_ACC[wname]


@pytest.mark.parametrize("N,n,m,wname", [((8, 16, 32), (16, 32, 64), 3, "gauss"),
                                         ((16, 8, 4), (32, 16, 8), 2, "kb"),
                                         ((12, 10, 6), (32, 16, 16), 5, "bspline"),
                                         ((4, 4, 4), (64, 64, 64), 8, "sincpow")])
def test_fp64_anisotropic(gpu, oracle, N, n, m, wname):
    say("case anisotropic", N, n, m, wname)
    rng = SplitMix64(11)
    M = 777
    x = rng.uniform(3 * M).reshape(M, 3) - 0.5
    fh = rng.complex_normal(int(np.prod(N)))
    fv = rng.complex_normal(M)
    g = run_gpu(N, n, m, WINDOWS[wname], x, fh, fv)
    P = oracle.Plan(N, n, m, WINDOWS[wname])
    f_o, g_o = P.trafo(x, fh, compute_grad=True)
    assert rel(g["f"], f_o) < 1e-12
    assert rel(g["grad"], g_o) < 1e-12
    assert rel(g["adj"], P.adj(x, f=fv)) < 1e-12


@pytest.mark.parametrize("wname", ["kb", "gauss"])
def test_fp32_vs_fp64_oracle(gpu, oracle, wname):
    w = make(1)
    say("case fp32", wname)
    x32 = w["x"].astype(np.float32).astype(np.float64)  # the FP32 plan sees float nodes
    g = run_gpu(w["N"], w["n"], w["m"], WINDOWS[wname], x32, w["f_hat"], w["f"], precision="float")
    P = oracle.Plan(w["N"], w["n"], w["m"], WINDOWS[wname])
    f_o, g_o = P.trafo(x32, w["f_hat"], compute_grad=True)
    assert rel(g["f"], f_o) < 1e-5
    assert rel(g["grad"], g_o) < 1e-5
    assert rel(g["adj"], P.adj(x32, f=w["f"])) < 1e-5


def test_edge_nodes_and_sizes(gpu, oracle):
    N, n, m = (8, 8, 8), (16, 16, 16), 3
    P = oracle.Plan(N, n, m, 0)
    rng = SplitMix64(12)
    special = np.array([[-0.5, -0.5, -0.5], [0.0, 0.0, 0.0], [0.25, -0.125, 0.4375],  # on grid points
                        [0.5 - 2 ** -40, -0.5 + 2 ** -40, 0.5 - 1e-9], [1 / 3, -1 / 7, 2 / 9]])
    for M in [1, 5, 257, 1000]:
        say("case edge M", M)
        x = np.concatenate([special, rng.uniform(3 * M).reshape(M, 3) - 0.5])[:M]
        fh = rng.complex_normal(512)
        fv = rng.complex_normal(M)
        g = run_gpu(N, n, m, 0, x, fh, fv)
        f_o, g_o = P.trafo(x, fh, compute_grad=True)
        assert rel(g["f"], f_o) < 1e-12
        assert rel(g["adj"], P.adj(x, f=fv)) < 1e-12
        # Kaiser-Bessel psi' = n x/d (psi - b cosh(b r)/pi) (kernel/ndft-parallel.c:2268) cancels as
        # d -> 0, i.e. when n x is within ~1e-6 of a grid line: there two libms differ by ~eps/d.
        nx = np.asarray(n) * x
        frac = nx - np.floor(nx)
        near = np.any((frac < 1e-6) & (frac > 0) | (frac > 1 - 1e-6), axis=1)
        g_gpu = g["grad"].reshape(-1, 3)
        g_ref = g_o.reshape(-1, 3)
        assert rel(g_gpu[~near], g_ref[~near]) < 1e-12
        if near.any():
            assert rel(g_gpu[near], g_ref[near]) < 1e-8


def test_empty_node_set(gpu):
    say("case M=0")
    dp = DevicePlan((8, 8, 8), (16, 16, 16), 3, 0, 0)
    dp.set_f_hat(np.ones(512))
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
    dp.adj(PA.PNFFT_COMPUTE_F)
    assert float(dp.f_hat.abs().max()) == 0.0  # adj of no nodes is zero
    assert dp.api.last_error() is None
    dp.close()


def test_accumulate_flags(gpu, oracle):
    w = make(1, M=300)
    say("case accumulate")
    dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"])
    dp.set_x(w["x"])
    dp.set_f_hat(w["f_hat"])
    dp.trafo(PA.PNFFT_COMPUTE_F)
    f1 = dp.f.cpu().numpy().copy()
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_ACCUMULATED)
    assert rel(dp.f.cpu().numpy(), 2 * f1) < 1e-14
    dp.set_f(w["f"])
    dp.adj(PA.PNFFT_COMPUTE_F)
    a1 = dp.f_hat.cpu().numpy().copy()
    dp.adj(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_ACCUMULATED)
    assert rel(dp.f_hat.cpu().numpy(), 2 * a1) < 1e-14
    dp.close()


def test_host_pointer_mode(gpu, oracle):
    """The reference's memory model: library-allocated host arrays written through the getters."""
    w = make(1, M=500)
    say("case host pointers")
    api = PA.lib().d
    plan = api.init_guru(3, api.arr3(w["N"]), api.arr3(w["n"]), (ctypes.c_double * 3)(0.5, 0.5, 0.5), w["m"],
                         PA.PNFFT_MALLOC_F_HAT, 0, PA.MPI_COMM_WORLD)
    nodes = api.init_nodes(w["M"], PA.PNFFT_MALLOC_X | PA.PNFFT_MALLOC_F | PA.PNFFT_MALLOC_GRAD_F)
    M, Nt = w["M"], int(np.prod(w["N"]))

    def view(ptr, count):
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)), shape=(count,))

    view(api.get_x(nodes), 3 * M)[:] = w["x"].reshape(-1)
    view(api.get_f_hat(plan), 2 * Nt)[:] = w["f_hat"].view(np.float64)
    api.trafo(plan, nodes, PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
    f = view(api.get_f(nodes), 2 * M).view(np.complex128).copy()
    gr = view(api.get_grad_f(nodes), 6 * M).view(np.complex128).copy()
    P = oracle.Plan(w["N"], w["n"], w["m"], 0)
    f_o, g_o = P.trafo(w["x"], w["f_hat"], compute_grad=True)
    assert rel(f, f_o) < 1e-12 and rel(gr, g_o) < 1e-12
    view(api.get_f(nodes), 2 * M)[:] = w["f"].view(np.float64)
    api.adj(plan, nodes, PA.PNFFT_COMPUTE_F)
    a = view(api.get_f_hat(plan), 2 * Nt).view(np.complex128).copy()
    assert rel(a, P.adj(w["x"], f=w["f"])) < 1e-12
    assert api.last_error() is None
    api.free_nodes(nodes, PA.PNFFT_FREE_X | PA.PNFFT_FREE_F | PA.PNFFT_FREE_GRAD_F)
    api.finalize(plan, PA.PNFFT_FREE_F_HAT)


def test_direct_ndft_matches_oracle(gpu, oracle):
    w = make(1, M=200)
    say("case direct")
    rng = SplitMix64(13)
    gv = rng.complex_normal(3 * w["M"])
    dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"])
    dp.set_x(w["x"])
    dp.set_f_hat(w["f_hat"])
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F | PA.PNFFT_COMPUTE_DIRECT)
    f_nd, g_nd = oracle.ndft_trafo(w["N"], w["x"], w["f_hat"], compute_grad=True)
    assert rel(dp.f.cpu().numpy(), f_nd) < 1e-12
    assert rel(dp.grad.cpu().numpy(), g_nd) < 1e-12
    dp.set_f(w["f"])
    dp.set_grad(gv)
    dp.adj(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F | PA.PNFFT_COMPUTE_DIRECT)
    assert rel(dp.f_hat.cpu().numpy(), oracle.ndft_adj(w["N"], w["x"], f=w["f"], grad=gv)) < 1e-12
    dp.close()


def _tile_keys(x, n, m, precision):
    # tiles of 8 cells per axis for m <= 4 (and FP64 m <= 6), of 4 cells otherwise; halved
    # on axes the tile size does not divide
    ts = []
    for nt in n:
        t = 8 if m <= 4 or (m <= 6 and precision == "double") else 4
        while t > 1 and nt % t:
            t //= 2
        ts.append(t)
    nt_ = [(a + b - 1) // b for a, b in zip(n, ts)]
    st = (np.floor(np.asarray(n, float) * x) - m).astype(np.int64) % np.asarray(n)
    tile = ((st[:, 0] // ts[0]) * nt_[1] + st[:, 1] // ts[1]) * nt_[2] + st[:, 2] // ts[2]
    # cell order inside a tile: runs of q = min(ts2, 4) cells along axis 2 outermost, then
    # axis 0, axis 1 and the cell within the run
    q = min(ts[2], 4)
    l0, l1, l2 = st[:, 0] % ts[0], st[:, 1] % ts[1], st[:, 2] % ts[2]
    inn = ((l2 // q) * (ts[0] * ts[1]) + l0 * ts[1] + l1) * q + l2 % q
    return (tile * (ts[0] * ts[1] * ts[2]) + inn).astype(np.int64)


@pytest.mark.parametrize("cfg,M,precision", [(1, 4096, "double"), (3, 200000, "float"), (5, 100001, "float")])
def test_binning_bit_exact(gpu, cfg, M, precision):
    w = make(cfg, M=M)
    say("case binning", cfg, M)
    x = w["x"].copy()
    x[::97] = x[5]  # ties keep input order
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    dp = DevicePlan(w["N"], w["n"], w["m"], w["flags"], M, precision=precision)
    dp.set_x(x)
    order, keys = dp.node_bins()
    order = order.cpu().numpy()
    keys = keys.cpu().numpy().view(np.uint32).astype(np.int64)
    k = _tile_keys(x, w["n"], w["m"], precision)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(order, ref)
    assert np.array_equal(keys, k[ref])
    dp.close()


def test_grad_x_vjp(gpu, oracle):
    w = make(1, M=600)
    say("case vjp")
    rng = SplitMix64(21)
    wv = rng.complex_normal(w["M"])
    dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"])
    dp.set_x(w["x"])
    dp.set_f_hat(w["f_hat"])
    gx = dp.grad_x_vjp(torch.as_tensor(wv).to("cuda")).cpu().numpy()
    _, g_o = oracle.Plan(w["N"], w["n"], w["m"], 0).trafo(w["x"], w["f_hat"], compute_grad=True)
    ref = (np.conj(wv)[:, None] * g_o.reshape(-1, 3)).real
    assert rel(gx, ref) < 1e-12
    dp.close()


def test_precompute_then_move_nodes(gpu, oracle):
    """Bins from precompute_psi are reused; set_x invalidates them."""
    w = make(1, M=1000)
    say("case precompute")
    P = oracle.Plan(w["N"], w["n"], w["m"], 0)
    dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"])
    dp.set_x(w["x"])
    dp.set_f_hat(w["f_hat"])
    dp.precompute()
    dp.trafo(PA.PNFFT_COMPUTE_F)
    assert rel(dp.f.cpu().numpy(), P.trafo(w["x"], w["f_hat"])[0]) < 1e-12
    x2 = np.roll(w["x"], 1, axis=1) * 0.9
    dp.set_x(x2)
    dp.trafo(PA.PNFFT_COMPUTE_F)
    assert rel(dp.f.cpu().numpy(), P.trafo(x2, w["f_hat"])[0]) < 1e-12
    dp.close()


@pytest.mark.parametrize("precision,tol,m", [("double", 1e-12, 6), ("float", 1e-5, 6), ("float", 1e-5, 4)])
def test_stale_bins_after_in_place_move(gpu, oracle, precision, tol, m):
    """precompute_psi bins reused after x changed in place (no set_x): nodes that left their
    tiles go through the direct paths, trafo (f and grad) and adj still match the oracle for the
    new x (FP32: the matrix-core gather and scatter, on tiles of 4 cells at m = 6 and of 8 at
    m = 4)."""
    w = make(1, M=1500)
    say("case stale bins", precision, m)
    P = oracle.Plan(w["N"], w["n"], m, 0)
    dp = DevicePlan(w["N"], w["n"], m, 0, w["M"], precision=precision)
    x = w["x"]
    if precision == "float":
        x = x.astype(np.float32).astype(np.float64)
    dp.set_x(x)
    dp.set_f_hat(w["f_hat"])
    dp.set_f(w["f"])
    dp.precompute()
    # move a third of the nodes by up to 3 cells, in place on the device
    rng = SplitMix64(77)
    x2 = x.copy()
    sel = np.arange(0, len(x), 3)
    x2[sel] = (x2[sel] + (rng.uniform(3 * len(sel)).reshape(-1, 3) - 0.5) * 6.0 / w["n"][0] + 0.5) % 1.0 - 0.5
    if precision == "float":
        x2 = x2.astype(np.float32).astype(np.float64)
    dp.x.copy_(torch.as_tensor(x2).to(dp.x.dtype))
    dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
    f_gpu = dp.f.cpu().numpy().copy()
    g_gpu = dp.grad.cpu().numpy().reshape(-1).copy()
    dp.set_f(w["f"])
    dp.adj(PA.PNFFT_COMPUTE_F)
    fh_gpu = dp.f_hat.cpu().numpy().copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    f_o, g_o = P.trafo(x2, w["f_hat"], compute_grad=True)
    assert rel(f_gpu, f_o) < tol
    assert rel(g_gpu, g_o) < tol
    assert rel(fh_gpu, P.adj(x2, f=w["f"])) < tol


@pytest.mark.parametrize("precision", ["double", "float"])
def test_adjoint_is_deterministic(gpu, precision):
    """The tile scatter folds fixed subgrid slots in a fixed order: repeated adj calls agree bit
    for bit (clustered nodes, many nodes per tile)."""
    w = make(3, M=1 << 16, scale=1 / 64)
    say("case deterministic adj", precision, w["N"], w["M"])
    dp = DevicePlan(w["N"], w["n"], w["m"], w["flags"], w["M"], precision=precision)
    dp.set_x(w["x"])
    dp.set_f(w["f"])
    outs = []
    for _ in range(3):
        dp.adj(PA.PNFFT_COMPUTE_F)
        outs.append(dp.f_hat.cpu().numpy().copy())
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("precision", ["double", "float"])
def test_precomputed_tables_equal_evaluated_windows(gpu, oracle, precision):
    """precompute_psi(PRE_PSI | PRE_GRAD_PSI) tables give the same trafo and grad bit for bit as
    evaluating the windows in the kernels (same code computes the tables), and the same adj up to
    summation order (the FP64 f+grad adjoint accumulates with atomics); both match the oracle."""
    w = make(1, M=3000)
    say("case precompute tables", precision)
    x = w["x"] if precision == "double" else w["x"].astype(np.float32).astype(np.float64)
    outs = []
    for pre in (0, PA.PNFFT_PRE_PSI | PA.PNFFT_PRE_GRAD_PSI):
        dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"], precision=precision)
        dp.set_x(x)
        dp.set_f_hat(w["f_hat"])
        if pre:
            dp.precompute(pre)
        dp.trafo(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
        f = dp.f.cpu().numpy().copy()
        g = dp.grad.cpu().numpy().copy()
        dp.set_f(w["f"])
        dp.set_grad(g)
        dp.adj(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
        fh = dp.f_hat.cpu().numpy().copy()
        err = dp.api.last_error()
        dp.close()
        assert err is None, err
        outs.append((f, g, fh))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert rel(outs[1][2], outs[0][2]) < (1e-14 if precision == "double" else 1e-7)
    tol = 1e-12 if precision == "double" else 1e-5
    P = oracle.Plan(w["N"], w["n"], w["m"], 0)
    f_o, g_o = P.trafo(x, w["f_hat"], compute_grad=True)
    assert rel(outs[1][0], f_o) < tol
    assert rel(outs[1][1], g_o) < tol
    assert rel(outs[1][2], P.adj(x, f=w["f"], grad=outs[1][1])) < tol


def test_fp32_adjoint_of_large_gradients(gpu, oracle):
    """FP32 adjoint of grad_f values from trafo (|g| ~ 1e4) with the Kaiser-Bessel window at m = 6,
    whose unscaled values reach 4e10 per axis: the kernels' power-of-two window scaling keeps the
    FP32 intermediates finite."""
    w = make(1, M=3000)
    say("case fp32 large gradients")
    x = w["x"].astype(np.float32).astype(np.float64)
    dp = DevicePlan(w["N"], w["n"], w["m"], 0, w["M"], precision="float")
    dp.set_x(x)
    dp.set_f_hat(w["f_hat"])
    dp.trafo(PA.PNFFT_COMPUTE_GRAD_F)
    g = dp.grad.cpu().numpy().copy()
    dp.set_f(w["f"])
    dp.adj(PA.PNFFT_COMPUTE_F | PA.PNFFT_COMPUTE_GRAD_F)
    fh = dp.f_hat.cpu().numpy().copy()
    err = dp.api.last_error()
    dp.close()
    assert err is None, err
    assert np.abs(g).max() > 1e3
    P = oracle.Plan(w["N"], w["n"], w["m"], 0)
    assert np.isfinite(fh).all()
    assert rel(fh, P.adj(x, f=w["f"], grad=g)) < 1e-5


@pytest.mark.parametrize("m,wname", [(2, "kb"), (3, "kb"), (4, "kb"), (5, "gauss"), (6, "kb"), (6, "bspline"),
                                     (8, "gauss")])
def test_fp32_modes_clustered(gpu, oracle, m, wname):
    """FP32 trafo and adjoint of f, of grad_f and of both (the matrix-core tile gather and
    scatter: tiles of 8 cells for m <= 4, of 4 cells for m = 5, 6; the scatter takes one, three
    and four rank-one terms per node; at m = 8 the vector gather and column scatter on 4-cell
    tiles) on clustered nodes: tiles from one node to several hundred (several batches per tile),
    with and without precomputed window tables."""
    from pnfft_amd.workloads import nodes
    rng = SplitMix64(400 + m + WINDOWS[wname])
    M = 20000
    N, n = (32, 32, 32), (64, 64, 64)
    x = nodes("gaussmix", M, rng).astype(np.float32).astype(np.float64)
    fh = rng.complex_normal(int(np.prod(N)))
    fv = rng.complex_normal(M)
    gv = rng.complex_normal(3 * M)
    P = oracle.Plan(N, n, m, WINDOWS[wname])
    F, G = PA.PNFFT_COMPUTE_F, PA.PNFFT_COMPUTE_GRAD_F
    f_o, g_o = P.trafo(x, fh, compute_grad=True)
    refs = {F: P.adj(x, f=fv), G: P.adj(x, grad=gv), F | G: P.adj(x, f=fv, grad=gv)}
    for pre in (0, PA.PNFFT_PRE_PSI | PA.PNFFT_PRE_GRAD_PSI):
        say("case fp32 modes m", m, wname, "pre", pre)
        dp = DevicePlan(N, n, m, WINDOWS[wname], M, precision="float")
        dp.set_x(x)
        if pre:
            dp.precompute(pre)
        errs = {}
        for flags in (F, G, F | G):
            dp.set_f_hat(fh)
            dp.trafo(flags)
            if flags & F:
                errs[("f", flags)] = rel(dp.f.cpu().numpy(), f_o)
            if flags & G:
                errs[("grad", flags)] = rel(dp.grad.cpu().numpy().reshape(-1), g_o)
        dp.set_f(fv)
        dp.set_grad(gv)
        for flags, ref in refs.items():
            dp.adj(flags)
            errs[("adj", flags)] = rel(dp.f_hat.cpu().numpy().reshape(-1), ref)
        err = dp.api.last_error()
        dp.close()
        assert err is None, err
        say("  rel errors", errs)
        assert max(errs.values()) < 1e-5


def _golden():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nfft_small.npz")
    z = np.load(path)  # allow_pickle stays False
    cases = sorted({k.split("__")[0] for k in z.files})
    return {c: {k.split("__")[1]: z[k] for k in z.files if k.startswith(c + "__")} for c in cases}


@pytest.mark.parametrize("case", ["kb_m3", "gauss_m3", "bspline_m3", "sincpow_m3", "kb_m6_aniso"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("float", 1e-5)])
def test_matches_golden_fixture(gpu, case, precision, tol):
    """trafo (f, grad) and adj (of f, of grad) against the committed FP64 fast-transform outputs of
    tests/golden/make_golden.py (every window; an anisotropic grid at m = 6)."""
    g = _golden()[case]
    say("case golden", case, precision)
    N, n, m, fl = [int(v) for v in g["N"]], [int(v) for v in g["n"]], int(g["m"]), int(g["flags"])
    x = g["x"] if precision == "double" else g["x"].astype(np.float32).astype(np.float64)
    out = run_gpu(N, n, m, fl, x, g["fhat"], g["f"], g["g"], precision=precision)
    assert rel(out["f"], g["nfft_f"]) < tol
    assert rel(out["grad"], g["nfft_grad"]) < tol
    assert rel(out["adj"], g["nfft_adj_f"]) < tol
    assert rel(out["adj_grad"], g["nfft_adj_grad"]) < tol
