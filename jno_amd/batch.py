"""Batched Jacobi-PCG: many systems of different sizes, or many right-hand sides on one operator,
in one device-driven solve (FP32 inner iterations, FP64 residual replacement, true-residual stop).

The reference batches with ``jax.vmap`` over the cg slot (LinearSolver traits ``vmap="native"``,
jno/utils/solver/solver_api.py:275-280; tests/test_fem_solver_slots.py:123-132), which runs every
member until the slowest converges and needs one operator shape.  Here every system keeps its own
recurrence and state on the device, so systems of different meshes share one launch sequence
without waiting for each other.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import torch

from . import _lib
from .fem import CSR
from .operator import SolverOperator


# launch shapes found by BatchSolver(autotune=True), per batch signature
_TUNED: dict = {}


class BatchSolver:
    """Batched mixed-precision Jacobi-PCG.

    The launch shape is per solver: ``one_kernel`` runs one kernel per step when every system is a
    Jacobi-scaled 4-diagonal DIA operator (3-D Kuhn lattices), else the three-kernel step;
    ``rows_per_entry`` (a multiple of 2048) and ``grid`` (a multiple of 8) shape the one-kernel
    step's table of (system, row range) entries, 0 taking the library defaults;
    ``steps_per_graph`` steps are captured in one hipGraph replay (0: direct launches).
    :func:`jno_amd.autotune.tune_batch` picks them by timing.  ``replace_factor``: the reduction of
    each system's FP32 residual norm between its FP64 residual replacements (0.003: four
    replacements to rtol 1e-10; on 256 uneven Kuhn-box systems 92.9 ms against 95.2 ms at 0.01 and
    103.4 ms at 0.1, the same iterations)."""

    def __init__(self, tol: float = 1e-10, atol: float = 0.0, maxiter: int = 20_000, replace_factor: float = 0.003,
                 steps_per_graph: int = 8, format: str = "auto", one_kernel: bool = True, rows_per_entry: int = 0,
                 grid: int = 0, autotune: bool = False):
        self.tol, self.atol, self.maxiter = float(tol), float(atol), int(maxiter)
        self.replace_factor = float(replace_factor)
        self.steps_per_graph = int(steps_per_graph)
        self.format = format
        self.tuning = _lib.JnoBatchTuning(1 if one_kernel else 0, int(rows_per_entry), int(grid), 0)
        # autotune: the first call on a batch signature (each system's size and layout) times the
        # launch shapes of autotune.tune_batch and keeps the fastest for later calls
        self.autotune = bool(autotune)
        self.last_launch_shape: dict = {}
        self.last_stats: List[dict] = []
        self.last_status = 0
        self.last_one_kernel = False  # whether the last call ran the one-kernel step

    def _op(self, A: CSR) -> SolverOperator:
        # the batch kernels stream DIA and CSR items only: the batch keeps its own cache entry, and
        # takes a single solver's cached layout (e.g. the DIA one of the fused assembly) only when it
        # is one of those, so neither solver's layout depends on which ran first
        key = ("batch_op", self.format)
        op = A._cache.get(key)
        if op is None:
            shared = A._cache.get(("solver_op", self.format, True))
            if shared is not None and shared.format in ("dia", "csr"):
                op = shared
            else:
                fmt, lanes = self.format, None
                if fmt == "auto":
                    from .autotune import choose_format

                    fmt, lanes = choose_format(A, allow=("dia", "csr"))
                elif fmt not in ("dia", "csr"):
                    raise ValueError("the batched solver streams 'dia' or 'csr' operators")
                op = SolverOperator(A, fmt=fmt, lanes=lanes, fp32=True)
            A._cache[key] = op
        return op

    def _signature(self, systems):
        sig = []
        for A, _ in systems:
            op = self._op(A)
            sig.append((A.shape[0], op.format, tuple(getattr(op, "offsets", ()) or ())))
        return (self.tol, self.atol, self.replace_factor, tuple(sig))

    def _tuned(self, systems):
        key = self._signature(systems)
        kw = _TUNED.get(key)
        if kw is None:
            from .autotune import tune_batch

            kw = tune_batch(list(systems), tol=self.tol, atol=self.atol, maxiter=self.maxiter,
                            replace_factor=self.replace_factor, format=self.format)
            _TUNED[key] = kw
        self.steps_per_graph = kw["steps_per_graph"]
        self.tuning.rows_per_entry = kw["rows_per_entry"]
        self.tuning.grid = kw["grid"]
        self.last_launch_shape = dict(kw)

    def __call__(self, systems: Sequence[Tuple[CSR, torch.Tensor]]) -> List[torch.Tensor]:
        """Solve A_i x_i = b_i for every (A_i, b_i); an operator object may repeat (shared)."""
        if not systems:
            return []
        if self.autotune:
            self._tuned(systems)
        items = (_lib.JnoBatchItem * len(systems))()
        xs, keep = [], []
        dev = systems[0][0].device
        for i, (A, b) in enumerate(systems):
            _lib.require_gpu(A.vals, f"A[{i}]")
            b = torch.as_tensor(b).reshape(-1).to(device=A.device, dtype=torch.float64).contiguous()
            if b.shape[0] != A.shape[0]:
                raise ValueError(f"system {i}: b has {b.shape[0]} entries for {A.shape[0]} rows")
            op = self._op(A)
            dinv = A.diag_inv()
            x = torch.empty_like(b)
            items[i].op = op.op
            items[i].dinv = dinv.data_ptr()
            items[i].b = b.data_ptr()
            items[i].x = x.data_ptr()
            xs.append(x)
            keep += [b, dinv, op]
        lib = _lib.load()
        wsb = lib.jno_batch_workspace_size(len(systems), items)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        stats = (_lib.JnoPcgStats * len(systems))()
        rc = lib.jno_batch_pcg_mixed_tuned(len(systems), items, self.tol, self.atol, self.maxiter,
                                           self.replace_factor, self.steps_per_graph, C.byref(self.tuning), stats,
                                           _lib.ptr(ws), wsb, _lib.stream_ptr(dev))
        self.last_status = int(rc)
        self.last_one_kernel = bool(self.tuning.ran_one_kernel)
        self.last_stats = [
            {"iterations": int(s.iterations), "refinements": int(s.refinements), "relres": float(s.relres),
             "seconds": float(s.seconds)} for s in stats
        ]
        if rc not in (_lib.JNO_OK, _lib.JNO_ENOTCONV, _lib.JNO_ENONFINITE):
            _lib.check(rc, "jno_batch_pcg_mixed")
        bad = [i for i, s in enumerate(self.last_stats) if not (s["relres"] <= 1e-4)]
        if bad:
            raise RuntimeError(
                f"jno.solve.cg did not solve {len(bad)} of {len(systems)} batched systems (first: {bad[0]}, "
                f"relative residual {self.last_stats[bad[0]]['relres']:.1e})"
            )
        return xs
