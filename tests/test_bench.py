"""Host logic of bench.py (no GPU): the one-GPU contract, the algorithmic byte and flop
formulas behind the roofline object, and the counter-calibration arithmetic of --pmc-traffic."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from pnfft_amd.workloads import CONFIGS  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_figures_of_the_headline_workload():
    w = dict(CONFIGS[3])
    ab = bench.algorithmic_bytes(w, "float")
    af = bench.algorithmic_flops(w)
    M, nt = w["M"], int(np.prod(w["n"]))
    # (2m+1)^3 stencil points, 4 flops each for f and 8 for the gradient: 2916 and 5832 per node at m = 4
    assert af["gather_f"] == 2916 * M and af["gather_grad"] == 5832 * M and af["scatter_f"] == 2916 * M
    # B stages: the grid once plus per node x (3 floats), its sort position and f (1 or 3 complex)
    assert ab["gather_f"] == 8 * nt + M * (12 + 4 + 8)
    assert ab["gather_grad"] == 8 * nt + M * (12 + 4 + 24)
    # the pruned FFT passes of one transform move c (N_tot + 2 N0 N1 n2 + 2 N0 n1 n2 + n_tot) bytes
    N0, N1, N2 = w["N"]
    n0, n1, n2 = w["n"]
    fwd = ab["fft_rows_fwd"] + ab["fft_cols1"] + ab["fft_cols0"]
    assert fwd == 8 * (N0 * N1 * N2 + 2 * N0 * N1 * n2 + 2 * N0 * n1 * n2 + n0 * n1 * n2)
    assert abs(fwd / 1e9 - 2.82) < 0.02


def _counter_csv(path, rows):
    with open(path, "w") as f:
        f.write("Kernel_Name,Counter_Name,Counter_Value\n")
        for name, value_kib in rows:
            f.write(f'"{name}",X,{value_kib}\n')


def test_traffic_calibration(tmp_path):
    gib_kib = (1 << 30) / 1024
    # the copy reads 1 GiB (FETCH counts half of it) and writes 1 GiB (WRITE counts all of it)
    _counter_csv(tmp_path / "cf.csv", [("__amd_rocclr_copyBuffer", gib_kib / 2)] * 3)
    _counter_csv(tmp_path / "cw.csv", [("__amd_rocclr_copyBuffer", gib_kib)] * 3)
    # two calls of each stage: two gather launches per call, raw 1 MiB fetched and 0.5 MiB written each
    g = "void pnfft_amd::(anonymous namespace)::gather_mfma_k<4, 1>(pnfft_amd::BArgs<float>)"
    _counter_csv(tmp_path / "f.csv", [(g, 1024.0)] * 4)
    _counter_csv(tmp_path / "w.csv", [(g, 512.0)] * 4)
    t = bench.pmc_traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                          str(tmp_path / "cw.csv"), "demo", steps_seen=2)
    assert abs(t["fetch_scale"] - 2.0) < 1e-9 and abs(t["write_scale"] - 1.0) < 1e-9
    c = t["classes"]["gather"]
    assert c["fetch_bytes_per_launch_raw"] == 1 << 20 and c["write_bytes_per_launch_raw"] == 1 << 19
    assert c["bytes_per_launch"] == 2 * (1 << 20) + (1 << 19)
    json.dumps(t)


def test_committed_traffic_file_matches_its_workload():
    name = CONFIGS[3]["name"]
    path = bench.traffic_file(name)
    t = json.load(open(path))
    assert t["workload"] == name
    assert set(t["classes"]) >= {"gather", "scatter", "fft_rows", "fft_cols"}
