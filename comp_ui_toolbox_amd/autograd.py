"""The operator pair as a differentiable PyTorch function.

BornApply.apply(op, gamma) returns u = A gamma; its backward is the
hand-written adjoint kernel (gamma_bar = A^H u_bar) of the same plan, so
losses such as 0.5 ||A gamma - u||^2 + lambda ||Psi gamma||_1 train through
the HIP kernels.  For complex tensors PyTorch's convention makes the
backward of a C-linear map the conjugate transpose applied to the incoming
gradient, which is exactly A^H; for a real gamma the gradient is its real
part.  BornBatchApply does the same for a batch of instances (one
persistent launch per direction).

The calls go through a binding of their own to the C library (the device
entry points born_forward_dev / born_adjoint_dev / born_batch_*_dev only),
separate from the one the tests and the host-buffer API use.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .operator_born import OperatorBorn, OperatorBatch

_dev = None


def _device_binding():
    """Second ctypes binding of libborn.so: device entry points only."""
    global _dev
    if _dev is None:
        _lib.lib()   # fails loudly when the library is missing
        L = C.CDLL(str(_lib.LIB_PATH))
        vp = C.c_void_p
        for name in ("born_forward_dev", "born_adjoint_dev"):
            f = getattr(L, name)
            f.argtypes = [vp, vp, C.c_int32, vp, vp]
            f.restype = C.c_int
        for name in ("born_batch_forward_dev", "born_batch_adjoint_dev"):
            f = getattr(L, name)
            f.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), vp]
            f.restype = C.c_int
        _dev = L
    return _dev


def _stream(t: torch.Tensor):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _apply(fn, plan, x: torch.Tensor, rows_out: int) -> torch.Tensor:
    x2 = x.reshape(1, -1) if x.dim() == 1 else x
    x2 = x2.to(torch.complex64).contiguous()
    out = torch.empty((x2.shape[0], rows_out), dtype=torch.complex64, device=x2.device)
    _lib.check(fn(plan, C.c_void_p(x2.data_ptr()), x2.shape[0], C.c_void_p(out.data_ptr()), _stream(x2)))
    return out.reshape(-1) if x.dim() == 1 else out


class BornApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op: OperatorBorn, gamma: torch.Tensor) -> torch.Tensor:
        ctx.op = op
        ctx.real_input = not gamma.is_complex()
        return _apply(_device_binding().born_forward_dev, op._plan, gamma, op.N_observations)

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        op = ctx.op
        g = _apply(_device_binding().born_adjoint_dev, op._plan, grad_u, op.N_points)
        if ctx.real_input:
            g = g.real
        return None, g


def born_apply(op: OperatorBorn, gamma: torch.Tensor) -> torch.Tensor:
    """u = A gamma with autograd support (gamma: (N_pts,) or (N_rhs, N_pts))."""
    return BornApply.apply(op, gamma)


def _batch_call(fn, batch: OperatorBatch, xs, rows_out):
    n = len(xs)
    xs = [x.to(torch.complex64).contiguous() for x in xs]
    outs = [torch.empty(r, dtype=torch.complex64, device=x.device) for x, r in zip(xs, rows_out)]
    pin = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
    pout = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
    _lib.check(fn(batch._batch, pin, pout, _stream(xs[0])))
    return outs


class BornBatchApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch: OperatorBatch, *gammas):
        ctx.batch = batch
        ctx.real = [not g.is_complex() for g in gammas]
        outs = _batch_call(_device_binding().born_batch_forward_dev, batch, gammas,
                           [op.N_observations for op in batch.ops])
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        batch = ctx.batch
        grads = [g if g is not None else torch.zeros(op.N_observations, dtype=torch.complex64, device="cuda")
                 for g, op in zip(grads, batch.ops)]
        gs = _batch_call(_device_binding().born_batch_adjoint_dev, batch, grads, [op.N_points for op in batch.ops])
        gs = [g.real if r else g for g, r in zip(gs, ctx.real)]
        return (None, *gs)


def born_batch_apply(batch: OperatorBatch, gammas):
    """[A_i gamma_i] for the instances of a batch, with autograd support."""
    return list(BornBatchApply.apply(batch, *gammas))
