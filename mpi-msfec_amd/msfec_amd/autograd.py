"""torch.autograd wrapper of the coarse misfit and its adjoint gradient.

``coarse_misfit(g, ...)`` returns J(g) = 1/2 |u_H(g) - u_target|^2 over the
free coarse nodes as a 0-d float64 tensor; its backward is the hand-written
adjoint of the basis solves and the coarse solve (msfec_q_adjoint_gradient),
not a traced graph.  g is the log-coefficient per fine cell (a = exp(g)),
shape (n_cells, N^dim), on the host.  The library does its own device work;
tensors cross the boundary as host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from .qbasis import QBasisBatch


class CoarseMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, vertices, dim, n_refine_local, nc, u_target, solver):
        gn = g.detach().cpu().numpy().astype(np.float64, copy=False)
        with QBasisBatch(vertices, dim=dim, n_refine_local=n_refine_local, solver=solver) as qb:
            qb.set_coefficient(np.exp(gn))
            qb.run()
            J, grad = qb.adjoint_gradient(nc, np.asarray(u_target, dtype=np.float64))
        ctx.save_for_backward(torch.from_numpy(grad).to(g.dtype))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None, None, None


def coarse_misfit(g, vertices, dim, n_refine_local, nc, u_target, solver=1):
    """J(g) with an adjoint backward; solver 1 = BAND_FP64_MFMA, 2 = BAND_MIXED."""
    return CoarseMisfit.apply(g, vertices, dim, n_refine_local, nc, u_target, solver)
