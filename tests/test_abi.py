"""The C interface: library loads and exports every declared symbol; the header
compiles as C and C++ and a C caller links against the library (CPU only, no
compute calls)."""
import os
import re
import subprocess

import pytest

import pnfft_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

# values of api/pnfft.h:302-418 in the reference
REFERENCE_FLAGS = {
    "PNFFT_PRE_PHI_HAT": 1 << 0, "PNFFT_FAST_GAUSSIAN": 1 << 1, "PNFFT_PRE_CONST_PSI": 1 << 2,
    "PNFFT_PRE_LIN_PSI": 1 << 3, "PNFFT_PRE_QUAD_PSI": 1 << 4, "PNFFT_PRE_CUB_PSI": 1 << 5,
    "PNFFT_MALLOC_F_HAT": 1 << 6, "PNFFT_FFT_OUT_OF_PLACE": 0, "PNFFT_FFT_IN_PLACE": 1 << 7,
    "PNFFT_INTERLACED": 1 << 8, "PNFFT_SHIFTED_F_HAT": 1 << 9, "PNFFT_SHIFTED_X": 1 << 10,
    "PNFFT_TRANSPOSED_NONE": 0, "PNFFT_TRANSPOSED_F_HAT": 1 << 11, "PNFFT_DIFF_AD": 0, "PNFFT_DIFF_IK": 1 << 12,
    "PNFFT_WINDOW_KAISER_BESSEL": 0, "PNFFT_WINDOW_GAUSSIAN": 1 << 13, "PNFFT_WINDOW_BSPLINE": 1 << 14,
    "PNFFT_WINDOW_SINC_POWER": 1 << 15, "PNFFT_WINDOW_BESSEL_I0": 1 << 16, "PNFFT_USE_FK_GAUSSIAN_T": 1 << 17,
    "PNFFT_WINDOW_GAUSSIAN_T": (1 << 17) | (1 << 13), "PNFFT_SORT_NODES": 1 << 18, "PNFFT_FREE_F_HAT": 1 << 6,
    "PNFFT_MALLOC_NONE": 0, "PNFFT_MALLOC_X": 1, "PNFFT_MALLOC_F": 2, "PNFFT_MALLOC_GRAD_F": 4,
    "PNFFT_MALLOC_HESSIAN_F": 8, "PNFFT_MALLOC_ALL": 15, "PNFFT_REAL_F": 1 << 4, "PNFFT_PRE_TENSOR": 0,
    "PNFFT_PRE_FULL": 1, "PNFFT_PRE_PSI": 2, "PNFFT_PRE_GRAD_PSI": 4, "PNFFT_PRE_HESSIAN_PSI": 8,
    "PNFFT_FREE_NONE": 0, "PNFFT_FREE_X": 1, "PNFFT_FREE_F": 2, "PNFFT_FREE_GRAD_F": 4,
    "PNFFT_FREE_HESSIAN_F": 8, "PNFFT_FREE_ALL": 15, "PNFFT_COMPUTE_F": 1, "PNFFT_COMPUTE_GRAD_F": 2,
    "PNFFT_COMPUTE_HESSIAN_F": 4, "PNFFT_COMPUTE_DIRECT": 8, "PNFFT_COMPUTE_ACCUMULATED": 16,
    "PNFFT_OMIT_DECONV": 32, "PNFFT_OMIT_FFT": 64, "PNFFT_OMIT_CONV": 128, "PNFFT_TIMER_ITER": 0,
    "PNFFT_TIMER_WHOLE": 1, "PNFFT_TIMER_LOOP_B": 2, "PNFFT_TIMER_SORT_NODES": 3, "PNFFT_TIMER_GCELLS": 4,
    "PNFFT_TIMER_MATRIX_B": 5, "PNFFT_TIMER_MATRIX_F": 6, "PNFFT_TIMER_MATRIX_D": 7,
    "PNFFT_TIMER_SHIFT_INPUT": 8, "PNFFT_TIMER_SHIFT_OUTPUT": 9, "PNFFT_TIMER_LENGTH": 10,
}


def test_library_exports_every_declared_symbol():
    L = pnfft_amd.lib()
    names = pnfft_amd.header_symbols()
    assert len(names) > 140
    missing = [n for n in names if not hasattr(L.cdll, n)]
    assert not missing, missing


def test_probe_library_exports_the_bandwidth_probe():
    """bench.py --bw-probe loads the stream kernels from their own library (loaded, not called here)."""
    import ctypes
    path = os.path.join(os.path.dirname(pnfft_amd.__file__), "libpnfft_probe.so")
    assert os.path.exists(path), "build() makes libpnfft_probe.so"
    assert hasattr(ctypes.CDLL(path), "pnfft_amd_hbm_probe")


def test_reference_entry_points_present():
    names = set(pnfft_amd.header_symbols())
    for base in ["init_3d", "init_guru", "init_nodes", "get_x", "get_f", "get_f_hat", "get_grad_f", "trafo", "adj",
                 "precompute_psi", "finalize", "free_nodes", "malloc", "free", "psi", "dpsi", "inv_phi_hat",
                 "local_size_3d", "get_timer_trafo"]:
        for p in ("pnfft_", "pnfftf_"):
            assert p + base in names


def _cc_available(cc):
    return subprocess.run(["which", cc], capture_output=True).returncode == 0


def test_flag_values_match_reference(tmp_path):
    if not _cc_available("gcc"):
        pytest.skip("gcc missing")
    src = tmp_path / "flags.c"
    body = "\n".join(f'  printf("{k} %u\\n", (unsigned)({k}));' for k in REFERENCE_FLAGS)
    src.write_text('#include <stdio.h>\n#include "pnfft.h"\nint main(void){\n' + body + "\n  return 0;\n}\n")
    exe = tmp_path / "flags"
    subprocess.run(["gcc", "-std=c99", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {l.split()[0]: int(l.split()[1]) for l in out.strip().splitlines()}
    assert got == REFERENCE_FLAGS


@pytest.mark.parametrize("with_complex", [False, True])
def test_c_caller_links(tmp_path, with_complex):
    """A C program written against the reference API compiles and links."""
    if not _cc_available("gcc"):
        pytest.skip("gcc missing")
    pre = "#include <complex.h>\n" if with_complex else ""
    src = tmp_path / "caller.c"
    src.write_text(pre + r'''
#include "pnfft.h"
int main(int argc, char **argv) {
  ptrdiff_t N[3] = {16, 16, 16}, n[3] = {32, 32, 32}, local_N[3], local_N_start[3];
  double lo[3], up[3], x_max[3] = {0.5, 0.5, 0.5};
  if (argc > 5) {  /* never executed: link check only */
    pnfft_init();
    pnfft_local_size_3d(N, MPI_COMM_WORLD, PNFFT_TRANSPOSED_NONE, local_N, local_N_start, lo, up);
    pnfft_plan p = pnfft_init_guru(3, N, n, x_max, 6, PNFFT_MALLOC_F_HAT, PFFT_ESTIMATE, MPI_COMM_WORLD);
    pnfft_nodes nd = pnfft_init_nodes(100, PNFFT_MALLOC_X | PNFFT_MALLOC_F | PNFFT_MALLOC_GRAD_F);
    pnfft_complex *f_hat = pnfft_get_f_hat(p);
    double *x = pnfft_get_x(nd);
    pnfft_init_x_3d(lo, up, 100, x);
    pnfft_init_f_hat_3d(N, local_N, local_N_start, PNFFT_TRANSPOSED_NONE, f_hat);
    pnfft_precompute_psi(p, nd, PNFFT_PRE_PSI);
    pnfft_trafo(p, nd, PNFFT_COMPUTE_F | PNFFT_COMPUTE_GRAD_F);
    pnfft_adj(p, nd, PNFFT_COMPUTE_F);
    pnfft_complex *f = pnfft_get_f(nd), *g = pnfft_get_grad_f(nd);
    (void)f; (void)g;
    double *t = pnfft_get_timer_trafo(p);
    pnfft_timer_free(t);
    pnfft_free_nodes(nd, PNFFT_FREE_X | PNFFT_FREE_F | PNFFT_FREE_GRAD_F);
    pnfft_finalize(p, PNFFT_FREE_F_HAT);
    pnfftf_plan pf = pnfftf_init_3d(N, MPI_COMM_WORLD);
    pnfftf_finalize(pf, PNFFT_FREE_F_HAT);
    pnfft_cleanup();
  }
  return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(pnfft_amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-I", INC, str(src), "-o", str(exe), "-L", libdir, "-lpnfft_amd",
                    f"-Wl,-rpath,{libdir}", "-Wl,--allow-shlib-undefined"], check=True)


def test_header_is_valid_cpp(tmp_path):
    if not _cc_available("g++"):
        pytest.skip("g++ missing")
    src = tmp_path / "h.cpp"
    src.write_text('#include "pnfft.h"\nint main(){ pnfft_plan p = 0; (void)p; return 0; }\n')
    subprocess.run(["g++", "-fsyntax-only", "-I", INC, str(src)], check=True)


def test_no_cpu_fallback_in_product():
    """The product package must not import the oracle or any CPU compute path."""
    pkg = os.path.join(ROOT, "pnfft_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", txt, re.M), fn
                assert "liboracle" not in txt, fn


def test_kernel_time_classes_match_library(tmp_path):
    """PNFFT_KT_COUNT of the header equals what the library's kernel_times reports (null plan: no copy)."""
    if not _cc_available("gcc"):
        pytest.skip("gcc missing")
    src = tmp_path / "kt.c"
    src.write_text('#include <stdio.h>\n#include "pnfft.h"\nint main(void){ printf("%d\\n", PNFFT_KT_COUNT); '
                   'return 0; }\n')
    exe = tmp_path / "kt"
    subprocess.run(["gcc", "-std=c99", "-I", INC, str(src), "-o", str(exe)], check=True)
    count = int(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    import pnfft_amd as PA
    from pnfft_amd.device import DevicePlan

    for prec in ("double", "float"):
        assert PA.lib().api(prec).kernel_times(None, None, None, 0) == count
    assert len(DevicePlan.KT_NAMES) == count


@pytest.mark.parametrize("prec,m,c2r", [("double", 1, False), ("double", 9, False), ("float", 12, False),
                                        ("double", 7, True), ("double", 8, True)])
def test_unsupported_plans_are_refused(prec, m, c2r):
    """Cutoffs outside the compiled 2..8 and double-precision c2r plans with m >= 7 give a NULL
    plan and a message (validated before any device call)."""
    import ctypes
    api = pnfft_amd.lib().api(prec)
    R = api.real
    init = api.init_guru_c2r if c2r else api.init_guru
    p = init(3, api.arr3((16, 16, 16)), api.arr3((32, 32, 32)), (R * 3)(0.5, 0.5, 0.5), m, 0, 0,
             pnfft_amd.MPI_COMM_WORLD)
    assert not p
    assert api.last_error() is not None
