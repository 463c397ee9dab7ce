"""torch.autograd harness over the exchange kernels (test harness, not the
driver): the forward kernel chain computes V_B = V[N] from beads 1 and P
(quadratic_bosonic_exchange.cpp:34-132) and the backward is the hand-written
adjoint kernel, -dV_B/dx for both beads (the exterior spring forces of
:159-215 without the interior links).  Tensors are float64 CUDA tensors of
shape (N, D); the kernels read them as SoA copies on torch's current stream.
"""
import torch

from . import XPlan


class ExchangeVB(torch.autograd.Function):
    """V_B(x1, xP) with its gradient from the adjoint kernel."""

    @staticmethod
    def forward(ctx, x1, xP, plan):
        if not (x1.is_cuda and xP.is_cuda) or x1.dtype != torch.float64 or xP.dtype != torch.float64:
            raise TypeError("ExchangeVB takes float64 CUDA tensors")
        if x1.shape != xP.shape or x1.dim() != 2 or x1.shape != (plan.natoms, plan.ndim):
            raise ValueError("x1 and xP must both have shape (%d, %d)" % (plan.natoms, plan.ndim))
        x1s = x1.detach().t().contiguous()
        xPs = xP.detach().t().contiguous()
        vb = torch.empty(1, dtype=torch.float64, device=x1.device)
        stream = torch.cuda.current_stream(x1.device).cuda_stream
        plan.forward(x1s.data_ptr(), xPs.data_ptr(), vb.data_ptr(), None, None, stream)
        ctx.plan = plan
        ctx.save_for_backward(x1s, xPs)
        return vb.reshape(())

    @staticmethod
    def backward(ctx, g):
        x1s, xPs = ctx.saved_tensors
        D, N = x1s.shape
        grad = torch.empty((2, D, N), dtype=torch.float64, device=x1s.device)
        stream = torch.cuda.current_stream(x1s.device).cuda_stream
        ctx.plan.vb_grad(x1s.data_ptr(), xPs.data_ptr(), grad.data_ptr(), stream)
        # the kernel returns -dV_B/dx (a force)
        dx1 = grad[0].t().neg() * g
        dxP = grad[1].t().neg() * g
        return dx1, dxP, None


def exchange_potential(x1, xP, plan):
    """V_B of beads 1 and P (differentiable in both)."""
    return ExchangeVB.apply(x1, xP, plan)


def make_plan(natoms, ndim, beta_t, spring_k, pbc=False, size=0.0):
    return XPlan(natoms, ndim, beta_t, spring_k, pbc, size)
