"""Differentiable H-matvec for PyTorch-ROCm.

HMatVec.forward runs y = H x on the MI355X engine; backward runs the adjoint
kernel pass grad_x = H^T grad_y (bw_hip_rmatvec), so crack-opening inversions
(loss(H w)) get exact gradients without forming H.  The reference has no
adjoint: the only H^T available there is the dense transpose of its blocks.
"""
from __future__ import annotations

import torch


class HMatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, hmat, precision="fp64", original=True):
        ctx.hmat, ctx.precision, ctx.original = hmat, precision, original
        if x.dim() == 1:
            return hmat.matvec(x.contiguous(), original=original, precision=precision)
        return hmat.matvec_batched(x, original=original, precision=precision)

    @staticmethod
    def backward(ctx, gy):
        h = ctx.hmat
        if gy.dim() == 1:
            gx = h.rmatvec(gy.contiguous(), original=ctx.original, precision=ctx.precision)
        else:  # many columns: one MFMA-tiled H^T call
            gx = h.rmatvec_batched(gy, original=ctx.original, precision=ctx.precision)
        return gx, None, None, None


def hmatvec(hmat, x, precision: str = "fp64", original: bool = True):
    """y = H x with autograd support (hmat: a HipHmat built with adjoint=True)."""
    if not hmat.adjoint:
        raise ValueError("HipHmat must be created with adjoint=True for gradients")
    return HMatVec.apply(x, hmat, precision, original)
