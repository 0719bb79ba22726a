"""Linear-solver slots: ``cg(...)`` mirrors ``jno.solve.cg`` (jno/solve.py:191-206).

``cg(tol=, atol=, maxiter=)(A, b, M=jacobi(), x0=None) -> x`` runs the HIP Jacobi-PCG with the
semantics of ``jax.scipy.sparse.linalg.cg`` (x0 = 0 unless given; stop when r.r <= max(tol^2 b.b,
atol^2) or at maxiter), then applies the eager residual guard of ``_maybe_residual_check``
(jno/utils/solver/solver_api.py:224-240): a RuntimeError when ||b - A x|| / ||b|| > 1e-4 or is not
finite.  ``precision="mixed"`` runs FP32 inner iterations with FP64 residual replacement and stops
on the true FP64 residual.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import _lib
from .fem import CSR
from .operator import SolverOperator
from .precond import Diagonal


class LinearSolver:
    """A configured solver: ``solver(A, b, *, M=None, x0=None) -> x`` (solver_api.py:243-297)."""

    def __init__(self, name: str, tol: float, atol: float, maxiter: Optional[int], precision: str = "fp64",
                 format: str = "auto", replace_factor: float = 0.1, guard: bool = True):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.name = name
        self.tol, self.atol = float(tol), float(atol)
        self.maxiter = 20_000 if maxiter is None else int(maxiter)
        self.precision = precision
        self.format = format
        self.replace_factor = float(replace_factor)
        self.guard = guard
        self.key = (name, self.tol, self.atol, self.maxiter, precision, format, self.replace_factor)
        self.last_stats: dict = {}

    def __eq__(self, other):
        return isinstance(other, LinearSolver) and other.key == self.key

    def __hash__(self):
        return hash(self.key)

    def __repr__(self):
        return f"jno_amd.solve.{self.name}(tol={self.tol}, precision={self.precision!r})"

    def operator(self, A: CSR) -> SolverOperator:
        key = ("solver_op", self.format, self.precision == "mixed")
        op = A._cache.get(key)
        if op is None:
            op = SolverOperator(A, fmt=self.format, fp32=self.precision == "mixed")
            A._cache[key] = op
        return op

    def __call__(self, A, b, *, M=None, x0=None):
        if not isinstance(A, CSR):
            raise TypeError("A must be an assembled jno_amd CSR operator (FemLinearSystem.evaluate)")
        _lib.require_gpu(A.vals, "A")
        n = A.shape[0]
        b = torch.as_tensor(b).reshape(-1)
        _lib.require_gpu(b, "b")
        b = b.to(torch.float64).contiguous()
        if b.shape[0] != n:
            raise ValueError(f"b has {b.shape[0]} entries for a {n}-row operator")
        if M is not None and hasattr(M, "materialize"):
            M = M.materialize(A)
        if M is None:
            dinv = torch.ones(n, dtype=torch.float64, device=A.device)
        elif isinstance(M, Diagonal):
            dinv = M.dinv.to(torch.float64).contiguous()
        else:
            raise TypeError(
                "M must be None, jno_amd.precond.jacobi() or a jno_amd.precond.Diagonal applier: the fused "
                "HIP PCG applies diagonal preconditioners only"
            )
        op = self.operator(A)
        x = torch.empty(n, dtype=torch.float64, device=A.device)
        use_x0 = 0
        if x0 is not None:
            x.copy_(torch.as_tensor(x0, dtype=torch.float64, device=A.device).reshape(-1))
            use_x0 = 1
        mixed = self.precision == "mixed"
        lib = _lib.load()
        wsb = lib.jno_pcg_workspace_size_op(C.byref(op.op), 1 if mixed else 0)
        ws = torch.empty(wsb, dtype=torch.uint8, device=A.device)
        stats = _lib.JnoPcgStats()
        stream = _lib.stream_ptr(A.device)
        if mixed:
            rc = lib.jno_pcg_mixed(C.byref(op.op), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x), use_x0, self.tol,
                                   self.atol, self.maxiter, self.replace_factor, C.byref(stats), _lib.ptr(ws), wsb, stream)
        else:
            rc = lib.jno_pcg_f64(C.byref(op.op), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x), use_x0, self.tol, self.atol,
                                 self.maxiter, C.byref(stats), _lib.ptr(ws), wsb, stream)
        self.last_stats = {
            "iterations": int(stats.iterations),
            "refinements": int(stats.refinements),
            "relres": float(stats.relres),
            "relres_recursive": float(stats.relres_recursive),
            "seconds": float(stats.seconds),
            "status": int(rc),
            "operator": op.describe(),
        }
        if rc not in (_lib.JNO_OK, _lib.JNO_ENOTCONV, _lib.JNO_ENONFINITE):
            _lib.check(rc, f"jno_amd.solve.{self.name}")
        rel = float(stats.relres)
        if self.guard and (not (rel == rel) or rel in (float("inf"),) or rel > 1e-4 or rc == _lib.JNO_ENONFINITE):
            raise RuntimeError(
                f"jno.solve.{self.name} did not solve the system (relative residual {rel:.1e}); the problem may be "
                "singular/ill-posed or need a preconditioner. Try jno.solve.lu(), a precond= spec, or your own solve_fn."
            )
        return x


def cg(*, tol: float = 1e-8, atol: float = 0.0, maxiter: Optional[int] = 20_000, precision: str = "fp64",
       format: str = "auto", replace_factor: float = 0.1) -> LinearSolver:
    """Jacobi-PCG slot: the analogue of ``jno.solve.cg(tol=, atol=, maxiter=)``."""
    return LinearSolver("cg", tol, atol, maxiter, precision=precision, format=format, replace_factor=replace_factor)


def set_chunk(chunk: int):
    """Iterations enqueued per host poll of the device-side convergence state."""
    _lib.call("jno_pcg_set_chunk", int(chunk))


def set_pacing(enable: bool):
    """Paced chunks for the one-kernel FP32 inner loops (small first chunks, later ones sized by
    the observed convergence rate); False = fixed chunks of ``set_chunk``."""
    _lib.call("jno_pcg_set_pacing", 1 if enable else 0)


def _prec_index(precision: str) -> int:
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return 1 if precision == "mixed" else 0


def set_tuning(precision: str, block: int, grid_cap: int, k1_vector: int, split: int = 0):
    """Launch shape of the PCG iteration kernels for ``precision`` ('fp64' iterations or the FP32
    inner iterations of 'mixed'): threads per workgroup, maximum workgroups, DIA tile variant, and
    two (``split=0``) or three (``split=1``: p materialized before a plain SpMV) kernels per
    iteration."""
    _lib.call("jno_pcg_set_tuning", _prec_index(precision), int(grid_cap), int(k1_vector), int(block), int(split))


def get_tuning(precision: str):
    """Current ``(block, grid_cap, k1_vector, split)`` for ``precision``."""
    cap, vec, blk, spl = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _lib.call("jno_pcg_get_tuning", _prec_index(precision), C.byref(cap), C.byref(vec), C.byref(blk), C.byref(spl))
    return int(blk.value), int(cap.value), int(vec.value), int(spl.value)
