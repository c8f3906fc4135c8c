"""torch.autograd.Function over the adjoint gradient, for torch.autograd.gradcheck.

``objective(knots, obj)`` returns the objective of a knot vector (a float64
host tensor) through ``mz_objective_grad``; its backward returns the adjoint
gradient that call computed, scaled by the incoming gradient.  The optimizers
and the batched evaluator call the C-ABI directly; this wrapper exists so
that torch's own finite-difference checker can be run against the adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import Objective


class ObjectiveFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, knots: torch.Tensor, obj: Objective) -> torch.Tensor:
        kv = np.ascontiguousarray(knots.detach().cpu().numpy(), dtype=np.float64)
        F, g = obj.objective_grad(kv)
        ctx.save_for_backward(torch.from_numpy(np.asarray(g, dtype=np.float64).copy()))
        return torch.tensor(float(F), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


def objective(knots: torch.Tensor, obj: Objective) -> torch.Tensor:
    """Differentiable objective of one knot vector (host float64 tensor)."""
    return ObjectiveFunction.apply(knots, obj)
