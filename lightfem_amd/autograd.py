"""``torch.autograd.Function`` over the gfx950 forward solve and its adjoint.

PyTorch is the gradient harness only: the forward is the device pipeline
(assembly, Dirichlet elimination, PCG) of :class:`lightfem_amd.adjoint.ElasticProblem`,
the backward its adjoint PCG solve plus the element-sensitivity kernel. Tensors
cross the boundary as float64 host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from .adjoint import ElasticProblem


class _Displacement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, young: torch.Tensor, problem: ElasticProblem) -> torch.Tensor:
        y = young.detach().to("cpu", torch.float64).numpy()
        u = problem.solve(y)
        ctx.problem = problem
        ctx.token = problem.token
        ctx.young = y.copy()
        return torch.from_numpy(u).to(dtype=young.dtype, device=young.device)

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        g = grad_u.detach().to("cpu", torch.float64).numpy()
        grad = ctx.problem.vjp(ctx.token, ctx.young, np.ascontiguousarray(g))
        return torch.from_numpy(grad).to(dtype=grad_u.dtype, device=grad_u.device), None


def displacement(young: torch.Tensor, problem: ElasticProblem) -> torch.Tensor:
    """u(E) for per-element moduli `young` (shape [n_elems]), differentiable in `young`."""
    return _Displacement.apply(young, problem)
