"""torch.autograd wrapper of the NFFT for test harnesses.

f = trafo(x, f_hat) with
  d/d f_hat:  grad_f_hat = adj(grad_f)                      (pnfft_adj, the adjoint NFFT)
  d/d x:      grad_x[j,t] = Re(conj(grad_f_j) d f_j/d x_t)   (pnfft_trafo_grad_x_vjp: the
              window-derivative gather of PNFFT_COMPUTE_GRAD_F fused with the contraction)
following PyTorch's convention for complex tensors (grad = dL/dRe + i dL/dIm).
Both backward products run in libpnfft_amd; nothing is computed on the CPU.
"""
import torch

from . import PNFFT_COMPUTE_F


class NfftTrafo(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, f_hat, plan):
        plan.set_x(x.detach())
        plan.set_f_hat(f_hat.detach())
        plan.trafo(PNFFT_COMPUTE_F)
        ctx.save_for_backward(x.detach(), f_hat.detach())
        ctx.plan = plan
        return plan.f.clone()

    @staticmethod
    def backward(ctx, gout):
        x, f_hat = ctx.saved_tensors
        plan = ctx.plan
        gout = gout.contiguous()
        grad_x = grad_fh = None
        plan.set_x(x)
        if ctx.needs_input_grad[1]:
            plan.set_f(gout)
            plan.adj(PNFFT_COMPUTE_F)
            grad_fh = plan.f_hat.clone().reshape(f_hat.shape)
        if ctx.needs_input_grad[0]:
            plan.set_f_hat(f_hat)
            grad_x = plan.grad_x_vjp(gout).to(x.dtype)
        return grad_x, grad_fh, None


def nfft(x, f_hat, plan):
    """f = NFFT(x, f_hat) on `plan` (a pnfft_amd.device.DevicePlan of matching sizes)."""
    return NfftTrafo.apply(x, f_hat, plan)
