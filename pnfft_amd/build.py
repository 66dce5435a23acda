"""Build libpnfft_amd.so (HIP kernels + C interface) for gfx950 with hipcc.

Every HIP translation unit is compiled with packed FP32 disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) and hardware
floating-point atomics enabled (``-munsafe-fp-atomics``).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpnfft_amd.so")
# the HBM bandwidth probe of bench.py --bw-probe, a library of its own
PROBE_LIB = os.path.join(HERE, "libpnfft_probe.so")
SOURCES = ["fft.hip", "nfft_b.hip", "ndft.hip", "api.cpp"]
# b_kernels.hip is compiled once per (precision, kind, mode): (name suffix, defines)
B_VARIANTS = [
    (f"{r}_{k}_{m}", [f"-DPNFFT_B_REAL={'float' if r == 'f32' else 'double'}", f"-DPNFFT_B_KIND={ki}",
                      f"-DPNFFT_B_MODE={mi}"])
    for r in ("f32", "f64")
    for k, ki, modes in (("gather", 0, ((1, "f"), (2, "grad"), (3, "fgrad"), (4, "vjp"))),
                         ("scatter", 1, ((1, "f"), (2, "grad"))))
    for mi, m in modes
]
# scatter_cols.hip is compiled once per (precision, mode)
COLS_VARIANTS = [
    (f"{r}_{m}", [f"-DPNFFT_B_REAL={'float' if r == 'f32' else 'double'}", f"-DPNFFT_B_MODE={mi}"])
    for r in ("f32", "f64")
    for mi, m in ((1, "f"), (2, "grad"), (3, "fgrad"))
]
# pre_psi.hip, hessian.hip and window_tab.hip are compiled once per precision
PRE_VARIANTS = [(r, [f"-DPNFFT_B_REAL={'float' if r == 'f32' else 'double'}"]) for r in ("f32", "f64")]
ARCH = os.environ.get("PNFFT_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _newest_input(src):
    hdrs = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    return max(os.path.getmtime(p) for p in hdrs + [os.path.join(CSRC, src), __file__])


def _compile(job, force=True):
    src, suffix, defines = job
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + suffix + ".o")
    if not force and os.path.exists(obj) and os.path.getmtime(obj) > _newest_input(src):
        return obj
    cmd = [HIPCC] + FLAGS + defines + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True, force=True):
    """Compile every translation unit (force=False skips objects newer than their inputs)."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = ([("b_kernels.hip", "_" + sfx, d) for sfx, d in B_VARIANTS] +
            [("scatter_cols.hip", "_" + sfx, d) for sfx, d in COLS_VARIANTS] +
            [("pre_psi.hip", "_" + sfx, d) for sfx, d in PRE_VARIANTS] +
            [("hessian.hip", "_" + sfx, d) for sfx, d in PRE_VARIANTS] +
            [("window_tab.hip", "_" + sfx, d) for sfx, d in PRE_VARIANTS] + [(s, "", []) for s in SOURCES])
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 4))) as ex:
        objs = list(ex.map(lambda j: _compile(j, force), jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    ptmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-shared", "-x", "hip", os.path.join(CSRC, "hbm_probe.hip"), "-o", ptmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for hbm_probe.hip:\n{r.stderr[-6000:]}")
    os.replace(ptmp, PROBE_LIB)
    if verbose:
        print(f"built {LIB} and {PROBE_LIB}")
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
    sys.exit(0)
