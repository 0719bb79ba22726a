"""Differentiable solve: u(k) = A(k)^-1 b with a hand-written adjoint for dL/dk.

Replaces JAX reverse mode through ``custom_linear_solve(symmetric=True)`` inside jax's cg and the
VJP chain BCOO data -> Dirichlet mask -> segment_sum -> jacfwd of the element residual
(jno/utils/solver/fem_native.py:2326-2345, jno/utils/solver/fem_utils.py:3332-3336; reached via
jno/trace/__init__.py:3655-3669): backward solves A lam = dL/du with the same PCG (A is symmetric),
then contracts dL/dk_e = -(P lam)_e^T (K_e/k_e) u_e per element on the device.
"""
from __future__ import annotations

from typing import Optional

import torch

from .fem import FemLinearSystem
from .precond import jacobi
from .solve import LinearSolver, cg


class _PoissonSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, system: FemLinearSystem, solver: LinearSolver):
        A, b = system.evaluate({system.param: k.detach()})
        u = solver(A, b, M=jacobi())
        ctx.save_for_backward(u)
        ctx.system, ctx.solver, ctx.A = system, solver, A
        return u

    @staticmethod
    def backward(ctx, gu):
        (u,) = ctx.saved_tensors
        system: FemLinearSystem = ctx.system
        rhs = gu.detach().to(torch.float64).contiguous().clone()
        if system.dmask is not None:
            rhs[system.dmask.bool()] = 0.0
        lam = ctx.solver(ctx.A, rhs, M=jacobi())
        return system.grad_k(u, lam), None, None


def solve(system: FemLinearSystem, k: torch.Tensor, solver: Optional[LinearSolver] = None) -> torch.Tensor:
    """Nodal solution u(k), differentiable with respect to the per-element coefficient ``k``."""
    solver = solver or cg(tol=1e-10)
    return _PoissonSolve.apply(k, system, solver)
