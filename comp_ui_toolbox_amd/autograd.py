"""The operator pair as a differentiable PyTorch function.

BornApply.apply(op, gamma) returns u = A gamma; its backward is the
hand-written adjoint kernel (gamma_bar = A^H u_bar) of the same plan, so
losses such as 0.5 ||A gamma - u||^2 + lambda ||Psi gamma||_1 train through
the HIP kernels.  For complex tensors PyTorch's convention makes the
backward of a C-linear map the conjugate transpose applied to the incoming
gradient, which is exactly A^H.
"""
from __future__ import annotations

import torch

from .operator_born import OperatorBorn


class BornApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op: OperatorBorn, gamma: torch.Tensor) -> torch.Tensor:
        ctx.op = op
        ctx.real_input = not gamma.is_complex()
        g = gamma if gamma.is_complex() else gamma.to(torch.complex64)
        return op.forward_dev(g.to(torch.complex64).contiguous())

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        g = ctx.op.adjoint_dev(grad_u.to(torch.complex64).contiguous())
        if ctx.real_input:
            g = g.real
        return None, g


def born_apply(op: OperatorBorn, gamma: torch.Tensor) -> torch.Tensor:
    """u = A gamma with autograd support (gamma: (N_pts,) or (N_rhs, N_pts))."""
    return BornApply.apply(op, gamma)
