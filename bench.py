"""Throughput of the NFFT hot path on one MI355X.

Workload (one "step"): pnfft_trafo(COMPUTE_F) + pnfft_adj(COMPUTE_F) +
pnfft_trafo(COMPUTE_GRAD_F) on one plan and node set whose arrays are already
resident in HBM (device pointers handed to pnfft_set_*).  Node binning and the
window tables of pnfft_precompute_psi(PNFFT_PRE_PSI | PNFFT_PRE_GRAD_PSI), and the
plan's D tables, are setup, timed separately (setup_ms).

Default workload: N = 256^3 modes, oversampled n = 512^3, Kaiser-Bessel m = 4,
M = 2^24 nodes from a mixture of 8 Gaussians (sigma = 0.05, wrapped), computed
by the single-precision plans (pnfftf_*) with FP64 D tables and FP64
node-to-cell indices; the FP64 plans (pnfft_*) run the same inputs to give the
accuracy and the FP64 speed beside it.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nonuniform pts/s (trafo+adj+grad), N=256^3 M=2^24, FP32-comp vs FP64 at err≤1e-5"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(w, precision, c2r=False):
    """Algorithmic HBM bytes of one launch of each kernel class (every array read or written once).

    FFT passes: elements in + elements out (the pruned sides hold N, the padded sides n points);
    the backward rows pass adds D^T onto f_hat (read + write).  B stages: the oversampled grid once
    plus per node x (3 reals), its sort position (4 B) and f (1 complex) or grad_f (3 complex).
    """
    cs = 8 if precision == "float" else 16
    rs = cs // 2
    nt = int(np.prod(w["n"]))
    M = w["M"]
    N0, N1, N2 = w["N"]
    n0, n1, n2 = w["n"]
    nfh = N0 * N1 * (N2 // 2 + 1 if c2r else N2)
    b1 = N0 * N1 * n2
    b2 = N0 * n1 * n2
    node = 3 * rs + 4
    return {
        "fft_rows_fwd": cs * (nfh + b1),
        "fft_rows_bwd": cs * (b1 + 2 * nfh),
        "fft_cols1": cs * (b1 + b2),
        "fft_cols0": cs * (b2 + nt),
        "gather_f": cs * nt + M * (node + cs),
        "gather_grad": cs * nt + M * (node + 3 * cs),
        "scatter_f": cs * nt + M * (node + cs),
    }


def roofline_from_kernels(kt, ab, steps):
    """Dominant kernel class of one step and its achieved HBM rate.

    kt: {class: (ms, launches)} over the timed steps (HIP events on the plan's stream).  One step
    launches: FFT rows 2 forward + 1 backward, cols1 3, cols0 3; gather stage (tiled + direct
    kernels) 2 (f, grad); scatter stage (tiled kernel, fold of the tile subgrids into the grid,
    direct kernel for stale bins) 1.
    """
    def ms(k):
        return kt[k][0] / steps

    classes = {
        "fft_rows": (ms("fft_rows"), 2 * ab["fft_rows_fwd"] + ab["fft_rows_bwd"], 3),
        "fft_cols1": (ms("fft_cols1"), 3 * ab["fft_cols1"], 3),
        "fft_cols0": (ms("fft_cols0"), 3 * ab["fft_cols0"], 3),
        "gather_tiled+direct": (ms("gather_tiled") + ms("gather_direct"), ab["gather_f"] + ab["gather_grad"], 2),
        "scatter_tiled+fold+direct": (ms("scatter_tiled") + ms("scatter_fold") + ms("scatter_direct"),
                                      ab["scatter_f"], 1),
    }
    dom = max(classes, key=lambda k: classes[k][0])
    t_ms, nbytes, nl = classes[dom]
    achieved = (nbytes / nl) / (t_ms / nl * 1e-3) / 1e9 if t_ms > 0 else 0.0
    per_kernel = {k: round(v[0], 3) for k, v in classes.items()}
    per_kernel.update({k: round(ms(k), 3) for k in ("gather_tiled", "gather_direct", "scatter_tiled", "scatter_fold",
                                                    "scatter_direct")})
    return dom, achieved, nbytes / nl, t_ms / nl, per_kernel


def run_plan(w, precision, steps, warmup, x_np):
    import torch

    import pnfft_amd as PA
    from pnfft_amd.device import DevicePlan

    dp = DevicePlan(w["N"], w["n"], w["m"], w["flags"], w["M"], precision=precision)
    dp.set_x(x_np)
    dp.set_f_hat(w["f_hat"])
    dp.set_f(w["f"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dp.precompute(PA.PNFFT_PRE_PSI | PA.PNFFT_PRE_GRAD_PSI)  # window tables of psi and psi' (setup)
    torch.cuda.synchronize()
    setup_ms = 1e3 * (time.perf_counter() - t0)
    f_in = dp.f.clone()
    # trafo reads f_hat from buffer A, adj writes buffer B (values stay bounded over the steps)
    import ctypes
    fh_b = torch.empty_like(dp.f_hat)
    pa, pb = ctypes.c_void_p(dp.f_hat.data_ptr()), ctypes.c_void_p(fh_b.data_ptr())

    def step():
        dp.api.set_f_hat(pa, dp.plan)
        dp.trafo(PA.PNFFT_COMPUTE_F)
        dp.api.set_f_hat(pb, dp.plan)
        dp.adj(PA.PNFFT_COMPUTE_F)
        dp.api.set_f_hat(pa, dp.plan)
        dp.trafo(PA.PNFFT_COMPUTE_GRAD_F)

    # outputs of one step for the accuracy comparison
    dp.trafo(PA.PNFFT_COMPUTE_F)
    out_f = dp.f.clone()
    dp.f.copy_(f_in)
    dp.api.set_f_hat(pb, dp.plan)
    dp.adj(PA.PNFFT_COMPUTE_F)
    out_fh = fh_b.clone()
    dp.api.set_f_hat(pa, dp.plan)
    dp.trafo(PA.PNFFT_COMPUTE_GRAD_F)
    out_g = dp.grad.clone()
    for i in range(warmup):
        step()
    torch.cuda.synchronize()
    dp.reset_timers()
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        log(f"  {precision} step {i + 1}/{steps}")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    tt, ta = dp.timers()
    kt = dp.kernel_times()
    err = dp.api.last_error()
    if err:
        raise RuntimeError(err.decode())
    res = dict(ms_per_step=1e3 * el / steps, setup_ms=setup_ms, timer_trafo=tt, timer_adj=ta, kernel_times=kt,
               out=(out_f.cpu().numpy(), out_fh.cpu().numpy().reshape(-1), out_g.cpu().numpy().reshape(-1)))
    dp.close()
    del dp
    torch.cuda.empty_cache()
    return res


def cpu_baseline(w, sample_M, threads):
    """FP64 oracle (C restatement with OpenMP + scipy FFT) on the host cores."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import oracle as O

    P = O.Plan(w["N"], w["n"], w["m"], w["flags"], workers=threads)
    xs = w["x"][:sample_M]
    fs = w["f"][:sample_M]
    t = {}
    t0 = time.perf_counter()
    g2 = P.F(P.D(w["f_hat"]))
    t["DF"] = time.perf_counter() - t0
    gpad = np.ascontiguousarray(P.ghost_exchange(g2))
    del g2
    import ctypes
    f = np.zeros(sample_M, np.complex128)
    gr = np.zeros(3 * sample_M, np.complex128)
    L = O.lib()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
    n3 = np.asarray(P.n, np.int64)
    no3 = np.asarray(P.no, np.int64)
    xsc = np.ascontiguousarray(xs.reshape(-1))
    t0 = time.perf_counter()
    L.orc_B_trafo(lp(n3), lp(no3), P.m, P.window, dp(P.b), sample_M, dp(xsc), dp(gpad.view(np.float64)), 0,
                  dp(f.view(np.float64)), None, 1)
    t["B"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.orc_B_trafo(lp(n3), lp(no3), P.m, P.window, dp(P.b), sample_M, dp(xsc), dp(gpad.view(np.float64)), 0,
                  None, dp(gr.view(np.float64)), 2)
    t["Bgrad"] = time.perf_counter() - t0
    del gpad
    fsc = np.ascontiguousarray(fs.astype(np.complex128).reshape(-1))
    gpad = np.zeros([v + 2 * P.m for v in P.no], np.complex128)
    t0 = time.perf_counter()
    L.orc_B_adj_blocks(lp(n3), lp(no3), P.m, P.window, dp(P.b), sample_M, dp(xsc), dp(fsc.view(np.float64)), None,
                       0, dp(gpad.view(np.float64)), 1)
    t["BT"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    fh = np.zeros(int(np.prod(P.N)), np.complex128)
    P.DT_accumulate(P.FH(P.ghost_reduce(gpad)), fh)  # ghost fold + F^H + D^T
    t["FDT"] = time.perf_counter() - t0
    del gpad, fh
    scale = w["M"] / sample_M
    # D+F and F^H+D^T are independent of M (the step has two trafos and one adj); the B stages scale with M
    total = 2 * t["DF"] + t["FDT"] + (t["B"] + t["Bgrad"] + t["BT"]) * scale
    return dict(value=w["M"] / total, unit="pts/s", cores=threads, kind="port",
                sample=(f"FP64 oracle (reference-algorithm restatement) on N={w['N'][0]}^3 n={w['n'][0]}^3 "
                        f"m={w['m']}: D+F once ({t['DF']:.2f}s), gather f and grad and adj on {sample_M} of "
                        f"{w['M']} nodes ({t['B']:.2f}+{t['Bgrad']:.2f}+{t['BT']:.2f}s) scaled by {scale:.0f} to "
                        f"the full node count, F^H+D^T once ({t['FDT']:.2f}s); {threads} threads"),
                seconds=t)


def copy_probe():
    """Plain device copy of known size: 3 copies of 2^30 bytes (reads 2^30, writes 2^30 each)."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_read_per_launch": 1 << 30, "copy_bytes_written_per_launch": 1 << 30,
                      "launches": 3, "GB_per_s": round(3 * 2 * (1 << 30) / el / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload's node count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is "
              "described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    import torch

    from pnfft_amd.workloads import make

    log(f"workload config {args.config}, scale {args.scale}")
    w = make(args.config, scale=args.scale)
    precision = w["precision"]
    x_np = w["x"].astype(np.float32).astype(np.float64) if precision == "float" else w["x"]
    w["x"] = x_np
    log(f"  M={w['M']} N={w['N']} n={w['n']} m={w['m']} precision={precision}")

    res = run_plan(w, precision, args.steps, args.warmup, x_np)
    log(f"  {precision}: {res['ms_per_step']:.3f} ms/step")
    acc = None
    fp64 = None
    if precision == "float" and not args.no_fp64:
        fp64 = run_plan(w, "double", max(1, args.steps // 2), 1, x_np)
        log(f"  double: {fp64['ms_per_step']:.3f} ms/step")
        names = ("f", "f_hat", "grad_f")
        acc = {nm: float(np.linalg.norm(a - b) / np.linalg.norm(b))
               for nm, a, b in zip(names, res["out"], fp64["out"])}
        log(f"  rel-l2 vs FP64: {acc}")

    steps = args.steps
    ab = algorithmic_bytes(w, precision)
    dom, achieved, lbytes, lms, per_kernel = roofline_from_kernels(res["kernel_times"], ab, steps)
    log(f"  dominant kernel {dom}: {lms:.3f} ms/launch, {lbytes / 1e9:.3f} GB/launch, {achieved:.0f} GB/s")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": None, "kernel": dom,
                "launch_ms": round(lms, 4), "algorithmic_bytes_per_launch": int(lbytes),
                "kernel_ms_per_step": per_kernel}
    cpu = None
    if not args.no_cpu:
        log("  cpu baseline ...")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1))
        cpu = cpu_baseline(w, min(args.cpu_sample, w["M"]), threads)
        cpu.pop("seconds", None)
    value = w["M"] / (res["ms_per_step"] * 1e-3)
    line = {
        "metric": METRIC, "value": value, "unit": "pts/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if precision == "float" else "f64", "data": "synthetic",
        "config": {"workload": w["name"], "N": list(w["N"]), "n": list(w["n"]), "m": w["m"], "M": w["M"],
                   "window": "kaiser_bessel" if w["flags"] == 0 else w["flags"], "setup_ms": round(res["setup_ms"], 2)},
        "accuracy_rel_l2_vs_fp64": acc,
        "fp64_ms_per_step": fp64["ms_per_step"] if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
