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
from dataclasses import dataclass, replace
from typing import Optional

import torch

from . import _lib
from .fem import CSR
from .operator import SolverOperator
from .precond import Diagonal


@dataclass(frozen=True)
class Tuning:
    """Launch tuning of a solve (the C ``jno_tuning``), passed per call: ``(block, grid_cap, k1_vector,
    split)`` for the FP64 iterations and for the FP32 inner iterations of mixed solves, iterations
    per host poll, and paced chunks.  Hashable by value, as the reference's solver specs are
    (jno/utils/solver/solver_api.py:261-308), so it can be part of a solver's key."""

    fp64: tuple = (1024, 512, 0, 1)
    mixed: tuple = (256, 768, 3, 5)
    chunk: int = 16
    pacing: bool = True

    @staticmethod
    def defaults() -> "Tuning":
        """The library defaults (what a solve without its own tuning runs)."""
        t = _lib.JnoTuning()
        _lib.call("jno_pcg_default_tuning", C.byref(t))
        shape = lambda s: (int(s.block), int(s.grid_cap), int(s.k1_vector), int(s.split))  # noqa: E731
        return Tuning(shape(t.iter[0]), shape(t.iter[1]), int(t.chunk), bool(t.pacing))

    def with_shape(self, precision: str, block: int, grid_cap: int, k1_vector: int, split: int = 0) -> "Tuning":
        shape = (int(block), int(grid_cap), int(k1_vector), int(split))
        return replace(self, **{"mixed" if _prec_index(precision) else "fp64": shape})

    def to_c(self) -> "_lib.JnoTuning":
        t = _lib.JnoTuning()
        for i, (blk, cap, vec, spl) in enumerate((self.fp64, self.mixed)):
            t.iter[i].block, t.iter[i].grid_cap, t.iter[i].k1_vector, t.iter[i].split = blk, cap, vec, spl
        t.chunk = int(self.chunk)
        t.pacing = 1 if self.pacing else 0
        _lib.call("jno_pcg_check_tuning", C.byref(t))
        return t


class LinearSolver:
    """A configured solver: ``solver(A, b, *, M=None, x0=None) -> x`` (solver_api.py:243-297)."""

    def __init__(self, name: str, tol: float, atol: float, maxiter: Optional[int], precision: str = "fp64",
                 format: str = "auto", replace_factor: float = 0.1, guard: bool = True,
                 tuning: Optional[Tuning] = None):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.name = name
        self.tol, self.atol = float(tol), float(atol)
        self.maxiter = 20_000 if maxiter is None else int(maxiter)
        self.precision = precision
        self.format = format
        self.replace_factor = float(replace_factor)
        self.guard = guard
        self.tuning = tuning
        self.key = (name, self.tol, self.atol, self.maxiter, precision, format, self.replace_factor, tuning)
        self.last_stats: dict = {}

    def __eq__(self, other):
        return isinstance(other, LinearSolver) and other.key == self.key

    def __hash__(self):
        return hash(self.key)

    def __repr__(self):
        return f"jno_amd.solve.{self.name}(tol={self.tol}, precision={self.precision!r})"

    def operator(self, A: CSR) -> SolverOperator:
        if getattr(A, "format", None) == "ebe":  # an element-by-element operator is its own layout
            return A
        key = ("solver_op", self.format, self.precision == "mixed")
        op = A._cache.get(key)
        if op is None:
            op = SolverOperator(A, fmt=self.format, fp32=self.precision == "mixed")
            A._cache[key] = op
        return op

    def __call__(self, A, b, *, M=None, x0=None):
        from .ebe import ElementOperator

        if not isinstance(A, (CSR, ElementOperator)):
            raise TypeError("A must be an assembled jno_amd CSR operator (FemLinearSystem.evaluate) or an "
                            "ElementOperator")
        _lib.require_gpu(A.rowptr if isinstance(A, CSR) else A.E, "A")
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
        # the tuning travels with this call: the solver's own, else the one the autotuner attached to
        # the operator, else the library defaults (tuning = NULL)
        opc = op.op
        tun = self.tuning if self.tuning is not None else getattr(op, "tuning", None)
        tc = None
        if tun is not None:
            tc = tun.to_c()
            opc = _lib.JnoOp.from_buffer_copy(op.op)
            opc.tuning = C.addressof(tc)
        wsb = lib.jno_pcg_workspace_size_op(C.byref(opc), 1 if mixed else 0)
        ws = torch.empty(wsb, dtype=torch.uint8, device=A.device)
        stats = _lib.JnoPcgStats()
        stream = _lib.stream_ptr(A.device)
        if mixed:
            rc = lib.jno_pcg_mixed(C.byref(opc), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x), use_x0, self.tol,
                                   self.atol, self.maxiter, self.replace_factor, C.byref(stats), _lib.ptr(ws), wsb, stream)
        else:
            rc = lib.jno_pcg_f64(C.byref(opc), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x), use_x0, self.tol, self.atol,
                                 self.maxiter, C.byref(stats), _lib.ptr(ws), wsb, stream)
        del tc
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
       format: str = "auto", replace_factor: float = 0.1, tuning: Optional[Tuning] = None) -> LinearSolver:
    """Jacobi-PCG slot: the analogue of ``jno.solve.cg(tol=, atol=, maxiter=)``; ``tuning`` (a
    :class:`Tuning`) fixes the launch shape of this solver's calls."""
    return LinearSolver("cg", tol, atol, maxiter, precision=precision, format=format, replace_factor=replace_factor,
                        tuning=tuning)


def set_chunk(chunk: int):
    """Default iterations enqueued per host poll of the device-side convergence state (solves
    without their own :class:`Tuning`)."""
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
    """Default launch shape of the PCG iteration kernels for ``precision`` ('fp64' iterations or the
    FP32 inner iterations of 'mixed'), for solves without their own :class:`Tuning`: threads per
    workgroup, maximum workgroups, DIA tile variant, and the iteration mode (``split``)."""
    _lib.call("jno_pcg_set_tuning", _prec_index(precision), int(grid_cap), int(k1_vector), int(block), int(split))


def get_tuning(precision: str):
    """Current default ``(block, grid_cap, k1_vector, split)`` for ``precision``."""
    cap, vec, blk, spl = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _lib.call("jno_pcg_get_tuning", _prec_index(precision), C.byref(cap), C.byref(vec), C.byref(blk), C.byref(spl))
    return int(blk.value), int(cap.value), int(vec.value), int(spl.value)
