"""A reference-style C program runs trafo and adj through libpnfft_amd.so (GPU).

The program is written the way the reference's own checks use the library
(tests/check_trafo.c:47-169): pnfft_init_3d defaults (m = 6, n = 2N), node and
f_hat arrays allocated by the library and written through the getters on the
host, pnfft_local_size_3d borders for the nodes, the fast transform compared
with PNFFT_COMPUTE_DIRECT on the same plan.  It exits non-zero when the fast
and direct results differ by more than the window's accuracy.
"""
import os
import subprocess

import pytest

import pnfft_amd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pnfft.h"

static double rel_err(const pnfft_complex *a, const pnfft_complex *b, ptrdiff_t n) {
  double num = 0, den = 0;
  for (ptrdiff_t i = 0; i < n; i++) {
    num += cabs(a[i] - b[i]) * cabs(a[i] - b[i]);
    den += cabs(b[i]) * cabs(b[i]);
  }
  return sqrt(num / den);
}

int main(void) {
  const ptrdiff_t N[3] = {16, 16, 16}, M = 1000;
  ptrdiff_t local_N[3], local_N_start[3];
  double lower[3], upper[3];
  pnfft_init();
  pnfft_local_size_3d(N, MPI_COMM_WORLD, PNFFT_TRANSPOSED_NONE, local_N, local_N_start, lower, upper);
  pnfft_plan plan = pnfft_init_3d(N, MPI_COMM_WORLD);
  if (!plan) return 2;
  if (pnfft_get_m(plan) != 6) return 3;
  ptrdiff_t n[3];
  pnfft_get_n(plan, n);
  if (n[0] != 32 || n[1] != 32 || n[2] != 32) return 4;
  pnfft_nodes nodes = pnfft_init_nodes(M, PNFFT_MALLOC_X | PNFFT_MALLOC_F | PNFFT_MALLOC_GRAD_F);
  pnfft_complex *f_hat = pnfft_get_f_hat(plan), *f = pnfft_get_f(nodes), *grad = pnfft_get_grad_f(nodes);
  double *x = pnfft_get_x(nodes);
  const ptrdiff_t Nt = local_N[0] * local_N[1] * local_N[2];
  pnfft_init_f_hat_3d(N, local_N, local_N_start, PNFFT_TRANSPOSED_NONE, f_hat);
  srand(0);
  pnfft_init_x_3d(lower, upper, M, x);

  pnfft_complex *f1 = pnfft_alloc_complex(M), *g1 = pnfft_alloc_complex(3 * M);
  pnfft_complex *fh1 = pnfft_alloc_complex(Nt), *fs = pnfft_alloc_complex(M);
  pnfft_trafo(plan, nodes, PNFFT_COMPUTE_F | PNFFT_COMPUTE_GRAD_F);
  memcpy(f1, f, sizeof(pnfft_complex) * M);
  memcpy(g1, grad, sizeof(pnfft_complex) * 3 * M);
  pnfft_trafo(plan, nodes, PNFFT_COMPUTE_F | PNFFT_COMPUTE_GRAD_F | PNFFT_COMPUTE_DIRECT);
  const double ef = rel_err(f1, f, M), eg = rel_err(g1, grad, 3 * M);

  pnfft_init_f(M, f);
  memcpy(fs, f, sizeof(pnfft_complex) * M);
  pnfft_adj(plan, nodes, PNFFT_COMPUTE_F);
  memcpy(fh1, f_hat, sizeof(pnfft_complex) * Nt);
  memcpy(f, fs, sizeof(pnfft_complex) * M);
  pnfft_adj(plan, nodes, PNFFT_COMPUTE_F | PNFFT_COMPUTE_DIRECT);
  const double ea = rel_err(fh1, f_hat, Nt);
  printf("fast vs direct: f %.3e grad_f %.3e adj %.3e\n", ef, eg, ea);

  pnfft_free(f1); pnfft_free(g1); pnfft_free(fh1); pnfft_free(fs);
  pnfft_finalize(plan, PNFFT_FREE_F_HAT);
  pnfft_free_nodes(nodes, PNFFT_FREE_X | PNFFT_FREE_F | PNFFT_FREE_GRAD_F);
  pnfft_cleanup();
  return (ef < 1e-9 && eg < 1e-9 && ea < 1e-9) ? 0 : 1;
}
"""


def test_reference_style_c_program_runs(gpu, tmp_path):
    src = tmp_path / "check_trafo_gpu.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "check_trafo_gpu"
    libdir = os.path.dirname(pnfft_amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lpnfft_amd", f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    print("case c program", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
