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


def algorithmic_flops(w):
    """Algorithmic flops of one launch of the B kernels (tensor-product evaluation, FMA = 2 flops):
    (2m+1)^3 stencil points per node, 4 flops each for f (complex value times real weight) and 8
    for the three gradient sums."""
    c3 = (2 * w["m"] + 1) ** 3
    M = w["M"]
    return {"gather_f": 4 * c3 * M, "gather_grad": 8 * c3 * M, "scatter_f": 4 * c3 * M}


def window_name(flags):
    """The window a plan's flags select (include/pnfft.h window bits; none set: Kaiser-Bessel)."""
    names = ((1 << 17, "gaussian_t"), (1 << 13, "gaussian"), (1 << 14, "bspline"), (1 << 15, "sinc_power"), (1 << 16, "bessel_i0"))
    for bit, name in names:
        if flags & bit:
            return name
    return "kaiser_bessel"


# MI355X peaks: HBM3E 8 TB/s; dense FP32 matrix-core rate 157.3 TFLOP/s (FP64 78.6)
PEAK_HBM_GBPS = 8000.0
PEAK_MFMA_TFLOPS = {"float": 157.3, "double": 78.6}


def roofline_from_kernels(kt, ab, af, steps, precision, m=4):
    """Dominant kernel class of one step, with its binding roofline.

    kt: {class: (ms, launches)} over the timed steps (HIP events on the plan's stream).  One step
    launches: FFT rows 2 forward + 1 backward, cols1 3, cols0 3 (HBM-bound: GB/s of algorithmic
    bytes); gather stage 2 (f, grad) and scatter stage 1 (subgrid kernel, fold, stale nodes), whose
    FP32 kernels run on the matrix cores for m <= 6 (TFLOP/s of algorithmic flops).
    """
    def ms(k):
        return kt[k][0] / steps

    # the matrix-core B kernels: FP32 plans with m <= 6 (padded tile subgrids of at most 16 cells)
    mfma = precision == "float" and m <= 6
    classes = {
        "fft_rows": (ms("fft_rows"), 2 * ab["fft_rows_fwd"] + ab["fft_rows_bwd"], 0, 3),
        "fft_cols1": (ms("fft_cols1"), 3 * ab["fft_cols1"], 0, 3),
        "fft_cols0": (ms("fft_cols0"), 3 * ab["fft_cols0"], 0, 3),
        "gather": (ms("gather_tiled") + ms("gather_direct"), ab["gather_f"] + ab["gather_grad"],
                   af["gather_f"] + af["gather_grad"], 2),
        "scatter": (ms("scatter_tiled") + ms("scatter_fold") + ms("scatter_direct"), ab["scatter_f"],
                    af["scatter_f"], 1),
    }
    dom = max(classes, key=lambda k: classes[k][0])
    t_ms, nbytes, nflops, nl = classes[dom]
    lsec = t_ms / nl * 1e-3
    gbps = (nbytes / nl) / lsec / 1e9 if t_ms > 0 else 0.0
    out = {"kernel": dom, "launch_ms": round(t_ms / nl, 4), "algorithmic_bytes_per_launch": int(nbytes / nl),
           "hbm_GBps": round(gbps, 1)}
    if nflops and mfma:
        tf = (nflops / nl) / lsec / 1e12 if t_ms > 0 else 0.0
        peak = PEAK_MFMA_TFLOPS[precision]
        out.update(bound="mfma", achieved=round(tf, 2), peak=peak, unit="TFLOP/s", frac=round(tf / peak, 4),
                   algorithmic_flops_per_launch=int(nflops / nl))
    else:
        out.update(bound="hbm", achieved=round(gbps, 1), peak=PEAK_HBM_GBPS, unit="GB/s",
                   frac=round(gbps / PEAK_HBM_GBPS, 4))
    per_kernel = {k: round(v[0], 3) for k, v in classes.items()}
    per_kernel.update({k: round(ms(k), 3) for k in ("gather_tiled", "gather_direct", "scatter_tiled", "scatter_fold",
                                                    "scatter_direct")})
    out["kernel_ms_per_step"] = per_kernel
    return out


# kernels of each class in a rocprofv3 counter collection
TRAFFIC_CLASSES = {
    "gather": ("gather_mfma_k", "gather_single_k", "gather_tiled_k", "gather_direct_k", "gather_stale_k"),
    "scatter": ("scatter_mfma_k", "scatter_cols_k", "fold_k", "fold8_k", "fold8b_k", "fold4_k", "scatter_stale_k",
                "scatter_tiled_k", "scatter_direct_k"),
    "fft_rows": ("fft_rows", "fft_rows512"),
    "fft_cols": ("fft_cols",),
}
# launches per step of each class (each class's kernels are summed per launch)
CLASS_LAUNCHES = {"gather": 2, "scatter": 1, "fft_rows": 3, "fft_cols": 6}


def _counter_sums(path):
    """{kernel short name: (sum of the counter in bytes, dispatches)} from a rocprofv3 counter CSV
    (the size counters are in KiB)."""
    import csv
    import glob
    if os.path.isdir(path):
        path = glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)[0]
    tot = {}
    for r in csv.DictReader(open(path)):
        name = r["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].split("<")[0]
        name = name.split("::")[-1].split()[-1]
        v, n = tot.get(name, (0.0, 0))
        tot[name] = (v + float(r["Counter_Value"]) * 1024.0, n + 1)
    return tot


def pmc_traffic(fetch, write, copy_fetch, copy_write, workload, steps_seen):
    """HBM bytes per launch of each kernel class from FETCH_SIZE / WRITE_SIZE runs, each scaled by
    what the same counter reads for the plain 1 GiB copy (copy_probe)."""
    cf, cw = _counter_sums(copy_fetch), _counter_sums(copy_write)
    big_f = max(cf.values(), key=lambda v: v[0] / v[1])
    big_w = max(cw.values(), key=lambda v: v[0] / v[1])
    f_scale = (1 << 30) / (big_f[0] / big_f[1])
    w_scale = (1 << 30) / (big_w[0] / big_w[1])
    fs, ws = _counter_sums(fetch), _counter_sums(write)
    out = {"workload": workload, "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
           "classes": {}}
    for cls, names in TRAFFIC_CLASSES.items():
        fb = sum(v[0] for k, v in fs.items() if k in names)
        wb = sum(v[0] for k, v in ws.items() if k in names)
        launches = CLASS_LAUNCHES[cls] * steps_seen
        out["classes"][cls] = {"fetch_bytes_per_launch_raw": int(fb / launches),
                               "write_bytes_per_launch_raw": int(wb / launches),
                               "bytes_per_launch": int((fb * f_scale + wb * w_scale) / launches)}
    return out


def traffic_file(workload):
    return os.path.join(ROOT, "profiles", "traffic_%s.json" % workload.replace("^", "e"))


def run_plan(w, precision, steps, warmup, x_np, tables=True):
    """tables=False: the node bins are made once but no window tables, so every B kernel
    evaluates the windows of its nodes (the window-recomputation mode)."""
    import torch

    import pnfft_amd as PA
    from pnfft_amd.device import DevicePlan

    dp = DevicePlan(w["N"], w["n"], w["m"], w["flags"], w["M"], precision=precision)
    dp.set_x(x_np)
    dp.set_f_hat(w["f_hat"])
    dp.set_f(w["f"])
    if not tables:
        os.environ["PNFFT_AMD_TABLES"] = "0"
    # setup: node bins and window tables of psi and psi'.  The first call also
    # allocates their buffers (setup_first_ms); the second one, which bins and
    # tabulates again in full, is the cost per step when the nodes move (setup_ms)
    setup = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dp.precompute(PA.PNFFT_PRE_PSI | PA.PNFFT_PRE_GRAD_PSI)
        torch.cuda.synchronize()
        setup.append(1e3 * (time.perf_counter() - t0))
    os.environ.pop("PNFFT_AMD_TABLES", None)
    setup_first_ms, setup_ms = setup
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
    res = dict(ms_per_step=1e3 * el / steps, setup_ms=setup_ms, setup_first_ms=setup_first_ms, timer_trafo=tt, timer_adj=ta, kernel_times=kt,
               out=(out_f.cpu().numpy(), out_fh.cpu().numpy().reshape(-1), out_g.cpu().numpy().reshape(-1)))
    dp.close()
    del dp
    torch.cuda.empty_cache()
    return res


def cpu_baseline(w, sample_M, threads):
    """FP64 oracle (C restatement with OpenMP + scipy FFT) on the host cores, built with the
    reference's release flags for this host (-O3 -march=native -ffast-math)."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import oracle as O
    L = O.lib_fast()

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
    return dict(value=w["M"] / total, unit="pts/s", cores=threads, nproc=os.cpu_count(), kind="port",
                sample=(f"FP64 oracle (reference-algorithm restatement) on N={w['N'][0]}^3 n={w['n'][0]}^3 "
                        f"m={w['m']}: D+F once ({t['DF']:.2f}s), gather f and grad and adj on {sample_M} of "
                        f"{w['M']} nodes ({t['B']:.2f}+{t['Bgrad']:.2f}+{t['BT']:.2f}s) scaled by {scale:.0f} to "
                        f"the full node count, F^H+D^T once ({t['FDT']:.2f}s); {threads} threads, "
                        f"oracle built -O3 -march=native -ffast-math"),
                seconds=t, f=f, grad=gr)


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


def bw_probe(reps=20):
    """Achievable HBM bandwidth, unprofiled: torch's device copy, sum (read only) and fill (write
    only) of a 1 GiB buffer, and the hand-written stream kernels of pnfft_amd/libpnfft_probe.so
    (copy, nontemporal copy, read, write, and copies with the access patterns of the axis-0
    column pass and of the 512-point rows pass); one warm-up call each, then `reps` calls timed
    with events."""
    import ctypes
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    acc = torch.empty((), dtype=torch.float32, device="cuda")
    ops = {"copy": (lambda: dst.copy_(src), 2), "read": (lambda: torch.sum(src, (0,), out=acc), 1),
           "write": (lambda: dst.fill_(2.0), 1)}
    out = {"bytes_per_call": 1 << 30, "reps": reps}
    for name, (fn, mult) in ops.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[f"{name}_GBps"] = round(mult * (1 << 30) / (ms * 1e-3) / 1e9, 1)
        out[f"{name}_ms"] = round(ms, 4)
    del src, dst
    torch.cuda.empty_cache()
    lib = ctypes.CDLL(os.path.join(ROOT, "pnfft_amd", "libpnfft_probe.so"))
    lib.pnfft_amd_hbm_probe.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.pnfft_amd_hbm_probe.restype = ctypes.c_int
    res = (ctypes.c_double * 6)()
    if lib.pnfft_amd_hbm_probe(res, reps) != 0:
        raise RuntimeError("HBM probe kernels failed")
    for k, name in enumerate(("kernel_copy", "kernel_copy_nt", "kernel_read", "kernel_write",
                              "kernel_copy_cols0_pattern", "kernel_copy_rows_pattern")):
        out[f"{name}_GBps"] = round(res[k], 1)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload's node count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-recompute", action="store_true",
                    help="skip the run with the windows evaluated in the B kernels (value_window_recompute null)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22,
                    help="nodes of the CPU baseline's B stages (a quarter of the headline workload)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (calibrates the HBM byte counters)")
    ap.add_argument("--bw-probe", action="store_true",
                    help="only time a 1 GiB device copy, read and fill (achievable HBM bandwidth, unprofiled)")
    ap.add_argument("--pmc-traffic", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="rocprofv3 FETCH_SIZE / WRITE_SIZE outputs of a bench run (--steps 1 --warmup 0: "
                         "two calls of each stage) and of --copy-probe; writes profiles/traffic_<workload>.json")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.bw_probe:
        bw_probe()
        return
    if args.pmc_traffic:
        from pnfft_amd.workloads import CONFIGS
        name = CONFIGS[args.config]["name"]
        t = pmc_traffic(*args.pmc_traffic, workload=name, steps_seen=2)
        with open(traffic_file(name), "w") as fo:
            json.dump(t, fo, indent=1)
        print(json.dumps(t, indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the slab decomposition "
              "over several GPUs is not built", file=sys.stderr)
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

    # the same steps with the windows evaluated in the B kernels instead of read from tables
    recompute = None
    if not args.no_recompute:
        recompute = run_plan(w, precision, args.steps, 1, x_np, tables=False)
        log(f"  {precision}, windows recomputed: {recompute['ms_per_step']:.3f} ms/step")

    steps = args.steps
    ab = algorithmic_bytes(w, precision)
    af = algorithmic_flops(w)
    rl = roofline_from_kernels(res["kernel_times"], ab, af, steps, precision, w["m"])
    log(f"  dominant kernel {rl['kernel']}: {rl['launch_ms']:.3f} ms/launch, {rl['achieved']} {rl['unit']}")
    traffic = None
    tf = traffic_file(w["name"])
    if os.path.exists(tf):
        tj = json.load(open(tf))
        cls = "fft_cols" if rl["kernel"].startswith("fft_cols") else rl["kernel"]
        if tj.get("workload") == w["name"] and cls in tj["classes"]:
            traffic = tj["classes"][cls]["bytes_per_launch"]
    roofline = {"bound": rl["bound"], "achieved": rl["achieved"], "peak": rl["peak"], "unit": rl["unit"],
                "frac": rl["frac"], "traffic": traffic}
    roofline.update({k: v for k, v in rl.items() if k not in roofline})
    cpu = None
    acc_oracle = None
    if not args.no_cpu:
        log("  cpu baseline ...")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1))
        sM = min(args.cpu_sample, w["M"])
        cpu = cpu_baseline(w, sM, threads)
        cpu.pop("seconds", None)
        # the GPU outputs of the same nodes against the FP64 oracle's (fast-math build)
        f_o, g_o = cpu.pop("f"), cpu.pop("grad")
        acc_oracle = {"f": float(np.linalg.norm(res["out"][0][:sM] - f_o) / np.linalg.norm(f_o)),
                      "grad_f": float(np.linalg.norm(res["out"][2][:3 * sM] - g_o) / np.linalg.norm(g_o)),
                      "nodes": sM}
        log(f"  rel-l2 vs FP64 oracle on {sM} nodes: {acc_oracle}")
    value = w["M"] / (res["ms_per_step"] * 1e-3)
    line = {
        "metric": METRIC if args.config == 3 else
        f"nonuniform pts/s (trafo+adj+grad), {w['name']}, {'FP32' if precision == 'float' else 'FP64'}",
        "value": value, "unit": "pts/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if precision == "float" else "f64", "data": "synthetic",
        "config": {"workload": w["name"], "N": list(w["N"]), "n": list(w["n"]), "m": w["m"], "M": w["M"],
                   "window": window_name(w["flags"]), "setup_ms": round(res["setup_ms"], 2),
                   "setup_first_ms": round(res["setup_first_ms"], 2)},
        "accuracy_rel_l2_vs_fp64": acc,
        "accuracy_rel_l2_vs_oracle": acc_oracle,
        # the same rate with the node binning and window tables redone for every step
        "value_incl_setup": w["M"] / ((res["ms_per_step"] + res["setup_ms"]) * 1e-3),
        # the rate with the windows evaluated in the B kernels (no window tables; bins made once)
        "value_window_recompute": w["M"] / (recompute["ms_per_step"] * 1e-3) if recompute else None,
        "fp64_ms_per_step": fp64["ms_per_step"] if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
