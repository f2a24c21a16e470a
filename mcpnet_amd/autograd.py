"""torch.autograd wrapper around the B-spline MI forward and its hand-written adjoint.

Used only for checking the adjoint (against finite differences and the FP64 oracle gradient):
the forward computes the FP32 MI matrix on the GPU and L = sum_n m_n MI(pi_n, pj_n); the
backward calls mcpg_bspline_mi_backward_f32, which recomputes the spline weights and joint
tables from X instead of storing them.
"""
from __future__ import annotations

import numpy as np
import torch


class _MaskedMILoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, handle, pi, pj, m, bins, order):
        mi = handle.bspline_mi(X.detach().contiguous(), bins=bins, order=order, precision="fp32")
        ii = torch.from_numpy(np.asarray(pi, dtype=np.int64)).to(X.device)
        jj = torch.from_numpy(np.asarray(pj, dtype=np.int64)).to(X.device)
        mm = torch.from_numpy(np.asarray(m, dtype=np.float64)).to(X.device)
        L = (mi[ii, jj].double() * mm).sum()
        ctx.save_for_backward(X.detach())
        ctx.args = (handle, pi, pj, m, bins, order)
        return L

    @staticmethod
    def backward(ctx, gL):
        (X,) = ctx.saved_tensors
        handle, pi, pj, m, bins, order = ctx.args
        dX, _ = handle.bspline_mi_backward(X.contiguous(), pi, pj, m, bins=bins, order=order)
        return (gL.to(dX.dtype) * dX), None, None, None, None, None, None


def masked_mi_loss(handle, X: torch.Tensor, pi, pj, m, bins: int = 10, order: int = 3) -> torch.Tensor:
    """L = sum_n m_n * MI(X)[pi_n, pj_n] with a GPU adjoint; X is a float32 tensor on the GPU."""
    return _MaskedMILoss.apply(X, handle, pi, pj, m, bins, order)
