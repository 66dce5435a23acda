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


# Charges and dipoles (modelled on tests/check_charge_dipole.c:70-238): f_hat is the
# adjoint of the charges (PNFFT_COMPUTE_F) plus the adjoint of the dipoles
# (PNFFT_COMPUTE_GRAD_F); the energy of a charge is f q, its force grad_f q; the
# energy of a dipole is grad_f . p, its force the Hessian times p.  The fast
# transforms are compared with PNFFT_COMPUTE_DIRECT on the same plan.  argv[1] holds
# the plan flags (interlacing, ik-differentiation, window); the program exits
# non-zero when the relative errors exceed argv[2].
CHARGE_DIPOLE = r"""
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pnfft.h"

static void perform(const ptrdiff_t *N, ptrdiff_t Mc, ptrdiff_t Md, int m, unsigned pflags, unsigned direct,
                    pnfft_complex *energy, pnfft_complex *force) {
  ptrdiff_t n[3] = {2 * N[0], 2 * N[1], 2 * N[2]};
  ptrdiff_t local_N[3], local_N_start[3];
  double x_max[3] = {0.5, 0.5, 0.5}, lower[3], upper[3];
  pnfft_local_size_guru(3, N, n, x_max, m, MPI_COMM_WORLD, PNFFT_TRANSPOSED_NONE, local_N, local_N_start,
                        lower, upper);
  pnfft_plan plan = pnfft_init_guru(3, N, n, x_max, m, PNFFT_MALLOC_F_HAT | pflags, PFFT_ESTIMATE,
                                    MPI_COMM_WORLD);
  if (!plan) { fprintf(stderr, "plan rejected: %s\n", pnfft_last_error()); exit(2); }
  pnfft_nodes charges = pnfft_init_nodes(Mc, PNFFT_MALLOC_X | PNFFT_MALLOC_F | PNFFT_MALLOC_GRAD_F);
  pnfft_nodes dipoles = pnfft_init_nodes(Md, PNFFT_MALLOC_X | PNFFT_MALLOC_GRAD_F | PNFFT_MALLOC_HESSIAN_F);
  pnfft_complex *q = pnfft_alloc_complex(Mc), *p = pnfft_alloc_complex(3 * Md);
  pnfft_complex *f_hat = pnfft_get_f_hat(plan);
  pnfft_complex *cf = pnfft_get_f(charges), *cg = pnfft_get_grad_f(charges);
  pnfft_complex *dg = pnfft_get_grad_f(dipoles), *dh = pnfft_get_hessian_f(dipoles);
  srand(0);
  pnfft_init_f(Mc, q);
  pnfft_init_f(3 * Md, p);
  srand(1);
  pnfft_init_x_3d_adv(lower, upper, x_max, Mc, pnfft_get_x(charges));
  pnfft_init_x_3d_adv(lower, upper, x_max, Md, pnfft_get_x(dipoles));
  memcpy(cf, q, sizeof(pnfft_complex) * Mc);
  memcpy(dg, p, sizeof(pnfft_complex) * 3 * Md);
  const ptrdiff_t Nt = local_N[0] * local_N[1] * local_N[2];
  pnfft_adj(plan, charges, direct | PNFFT_COMPUTE_F);
  pnfft_complex *buf = pnfft_alloc_complex(Nt);
  memcpy(buf, f_hat, sizeof(pnfft_complex) * Nt);
  pnfft_adj(plan, dipoles, direct | PNFFT_COMPUTE_GRAD_F);
  for (ptrdiff_t k = 0; k < Nt; k++) f_hat[k] += buf[k];
  pnfft_trafo(plan, charges, direct | PNFFT_COMPUTE_F | PNFFT_COMPUTE_GRAD_F);
  for (ptrdiff_t j = 0; j < Mc; j++) {
    energy[j] = cf[j] * q[j];
    for (int t = 0; t < 3; t++) force[3 * j + t] = cg[3 * j + t] * q[j];
  }
  pnfft_trafo(plan, dipoles, direct | PNFFT_COMPUTE_GRAD_F | PNFFT_COMPUTE_HESSIAN_F);
  for (ptrdiff_t j = 0; j < Md; j++) {
    const pnfft_complex *v = dg + 3 * j, *h = dh + 6 * j, *pj = p + 3 * j;
    pnfft_complex *o = force + 3 * (Mc + j);
    energy[Mc + j] = v[0] * pj[0] + v[1] * pj[1] + v[2] * pj[2];
    o[0] = h[0] * pj[0] + h[1] * pj[1] + h[2] * pj[2];
    o[1] = h[1] * pj[0] + h[3] * pj[1] + h[4] * pj[2];
    o[2] = h[2] * pj[0] + h[4] * pj[1] + h[5] * pj[2];
  }
  if (pnfft_last_error()) { fprintf(stderr, "error: %s\n", pnfft_last_error()); exit(3); }
  pnfft_free(buf); pnfft_free(q); pnfft_free(p);
  pnfft_finalize(plan, PNFFT_FREE_F_HAT);
  pnfft_free_nodes(charges, PNFFT_FREE_ALL);
  pnfft_free_nodes(dipoles, PNFFT_FREE_ALL);
}

static double rel_err(const pnfft_complex *a, const pnfft_complex *b, ptrdiff_t n) {
  double num = 0, den = 0;
  for (ptrdiff_t i = 0; i < n; i++) {
    num += cabs(a[i] - b[i]) * cabs(a[i] - b[i]);
    den += cabs(b[i]) * cabs(b[i]);
  }
  return sqrt(num / den);
}

int main(int argc, char **argv) {
  const unsigned pflags = argc > 1 ? (unsigned)strtoul(argv[1], NULL, 0) : 0;
  const double tol = argc > 2 ? atof(argv[2]) : 1e-6;
  const ptrdiff_t N[3] = {16, 16, 16}, Mc = 300, Md = 300, M = Mc + Md;
  pnfft_init();
  pnfft_complex *e1 = pnfft_alloc_complex(M), *e2 = pnfft_alloc_complex(M);
  pnfft_complex *f1 = pnfft_alloc_complex(3 * M), *f2 = pnfft_alloc_complex(3 * M);
  perform(N, Mc, Md, 6, pflags, 0, e1, f1);
  perform(N, Mc, Md, 6, pflags, PNFFT_COMPUTE_DIRECT, e2, f2);
  const double ee = rel_err(e1, e2, M), ef = rel_err(f1, f2, 3 * M);
  printf("flags %#x: energy rel err %.3e, force rel err %.3e\n", pflags, ee, ef);
  pnfft_free(e1); pnfft_free(e2); pnfft_free(f1); pnfft_free(f2);
  pnfft_cleanup();
  return (ee < tol && ef < tol) ? 0 : 1;
}
"""


# tolerances: the window's accuracy at m = 6 for the Hessian (Kaiser-Bessel by window derivatives
# about 2e-9, ik about 2e-11; Gaussian about 2e-5), from the FP64 oracle at the same sizes
@pytest.mark.parametrize("flags,tol", [(0, 2e-8), (pnfft_amd.PNFFT_INTERLACED, 2e-8), (pnfft_amd.PNFFT_DIFF_IK, 1e-9),
                                       (pnfft_amd.PNFFT_DIFF_IK | pnfft_amd.PNFFT_INTERLACED, 1e-9),
                                       (pnfft_amd.PNFFT_WINDOW_GAUSSIAN, 2e-4)])
def test_charge_dipole_program(gpu, tmp_path, flags, tol):
    src = tmp_path / "check_charge_dipole_gpu.c"
    src.write_text(CHARGE_DIPOLE)
    exe = tmp_path / "check_charge_dipole_gpu"
    libdir = os.path.dirname(pnfft_amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lpnfft_amd", f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    print("case charge-dipole program", hex(flags), flush=True)
    r = subprocess.run([str(exe), hex(flags), str(tol)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
