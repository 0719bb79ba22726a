"""Preconditioner specs: ``jacobi()`` mirrors ``jno.precond.jacobi()`` (jno/precond.py:98-112, :221).

A spec is materialized against the operator: the inverse diagonal 1/d with the |d| > 1e-30 guard
(zero diagonals left unscaled).  The fused PCG kernels apply it as z = dinv * r.
"""
from __future__ import annotations

import torch


class Diagonal:
    """A diagonal preconditioner applier v -> dinv * v (also usable as the ``M=`` argument)."""

    def __init__(self, dinv: torch.Tensor):
        self.dinv = dinv

    def __call__(self, v):
        return self.dinv * v


class _Jacobi:
    key = ("jacobi",)

    def materialize(self, A) -> Diagonal:
        return Diagonal(A.diag_inv())

    def __eq__(self, other):
        return isinstance(other, _Jacobi)

    def __hash__(self):
        return hash(self.key)

    def __repr__(self):
        return "jno_amd.precond.jacobi()"


def jacobi() -> _Jacobi:
    return _Jacobi()
