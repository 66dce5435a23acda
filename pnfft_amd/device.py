"""A plan and a node set bound to torch device tensors through the C interface.

Used by the GPU tests, the smoke check and the benchmark: the arrays are
hipMalloc'd by torch and handed to ``pnfft_set_x`` / ``pnfft_set_f`` /
``pnfft_set_grad_f`` / ``pnfft_set_f_hat``, so the library works on them in
place (device-resident mode).  Every computation goes through
``libpnfft_amd.so``; nothing here computes on the CPU.
"""
import ctypes

import torch

from . import MPI_COMM_WORLD, lib

_CDT = {"double": torch.complex128, "float": torch.complex64}
_RDT = {"double": torch.float64, "float": torch.float32}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class DevicePlan:
    def __init__(self, N, n=None, m=6, flags=0, M=0, precision="double", x_max=(0.5, 0.5, 0.5),
                 device="cuda:0", alloc_grad=True, c2r=False):
        self.precision = precision
        self.api = lib().api(precision)
        self.N = tuple(int(v) for v in N)
        self.n = tuple(2 * v for v in self.N) if n is None else tuple(int(v) for v in n)
        self.m = int(m)
        self.flags = int(flags)
        self.M = int(M)
        self.c2r = bool(c2r)
        self.device = torch.device(device)
        R = self.api.real
        init = self.api.init_guru_c2r if self.c2r else self.api.init_guru
        self.plan = init(3, self.api.arr3(self.N), self.api.arr3(self.n), (R * 3)(*x_max), self.m, self.flags, 0,
                         MPI_COMM_WORLD)
        if not self.plan:
            err = self.api.last_error()
            raise RuntimeError("plan rejected: " + (err.decode() if err else "unknown"))
        self.nodes = self.api.init_nodes(self.M, 0)
        cdt, rdt = _CDT[precision], _RDT[precision]
        vdt = rdt if self.c2r else cdt  # c2r plans: real f and grad_f
        self.fh_shape = (self.N[0], self.N[1], self.N[2] // 2 + 1) if self.c2r else self.N
        self.x = torch.zeros((self.M, 3), dtype=rdt, device=self.device)
        self.f = torch.zeros(self.M, dtype=vdt, device=self.device)
        self.grad = torch.zeros((self.M, 3), dtype=vdt, device=self.device) if alloc_grad else None
        self.hess = None  # allocated by alloc_hessian()
        self.f_hat = torch.zeros(self.fh_shape, dtype=cdt, device=self.device)
        self.api.set_x(_ptr(self.x), self.nodes)
        self.api.set_f(_ptr(self.f), self.nodes)
        if self.grad is not None:
            self.api.set_grad_f(_ptr(self.grad), self.nodes)
        self.api.set_f_hat(_ptr(self.f_hat), self.plan)

    # inputs -------------------------------------------------------------------
    def set_x(self, x):
        self.x.copy_(torch.as_tensor(x).reshape(self.M, 3).to(self.x.dtype))
        self.api.set_x(_ptr(self.x), self.nodes)  # marks the bins stale

    def set_f_hat(self, fh):
        self.f_hat.copy_(torch.as_tensor(fh).reshape(self.fh_shape).to(self.f_hat.dtype))

    def set_f(self, f):
        self.f.copy_(torch.as_tensor(f).reshape(self.M).to(self.f.dtype))

    def set_grad(self, g):
        self.grad.copy_(torch.as_tensor(g).reshape(self.M, 3).to(self.grad.dtype))

    def alloc_hessian(self):
        """hessian_f: (M, 6) components xx, xy, xz, yy, yz, zz (real for c2r plans)."""
        self.hess = torch.zeros((self.M, 6), dtype=self.f.dtype, device=self.device)
        self.api.set_hessian_f(_ptr(self.hess), self.nodes)
        return self.hess

    # operations ---------------------------------------------------------------
    def _check(self):
        err = self.api.last_error()
        if err:
            raise RuntimeError(err.decode())

    def precompute(self, flags=None):
        torch.cuda.synchronize(self.device)
        from . import PNFFT_PRE_PSI
        self.api.precompute_psi(self.plan, self.nodes, PNFFT_PRE_PSI if flags is None else flags)

    def trafo(self, compute_flags=1):
        torch.cuda.synchronize(self.device)
        self.api.trafo(self.plan, self.nodes, compute_flags)

    def adj(self, compute_flags=1):
        torch.cuda.synchronize(self.device)
        self.api.adj(self.plan, self.nodes, compute_flags)

    def grad_x_vjp(self, w):
        torch.cuda.synchronize(self.device)
        w = torch.as_tensor(w).to(self.f.dtype).reshape(self.M).contiguous()
        gx = torch.empty((self.M, 3), dtype=self.x.dtype, device=self.device)
        self.api.trafo_grad_x_vjp(self.plan, self.nodes, _ptr(w), _ptr(gx))
        return gx

    def node_bins(self):
        order = torch.empty(self.M, dtype=torch.int32, device=self.device)
        keys = torch.empty(self.M, dtype=torch.int32, device=self.device)
        torch.cuda.synchronize(self.device)
        rc = self.api.node_bins(self.plan, self.nodes, _ptr(order), _ptr(keys))
        if rc != 0:
            self._check()
        return order, keys

    def timers(self):
        t = self.api.get_timer_trafo(self.plan)
        a = self.api.get_timer_adj(self.plan)
        out = ([t[i] for i in range(10)], [a[i] for i in range(10)])
        self.api.timer_free(t)
        self.api.timer_free(a)
        return out

    KT_NAMES = ("fft_rows", "fft_cols1", "fft_cols0", "gather_tiled", "gather_direct", "scatter_tiled",
                "scatter_direct", "binning", "scatter_fold")

    def kernel_times(self):
        """{kernel class: (device ms accumulated, launches)} since the plan was made or reset."""
        n = len(self.KT_NAMES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        self.api.kernel_times(self.plan, ms, cnt, n)
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KT_NAMES)}

    def reset_timers(self):
        self.api.reset_timer(self.plan)

    def close(self):
        if self.nodes:
            self.api.free_nodes(self.nodes, 0)
            self.nodes = None
        if self.plan:
            self.api.finalize(self.plan, 0)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
