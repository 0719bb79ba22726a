"""Autotuner for the solver-operator layout: DIA when it applies, else BSR3 or CSR with the number
of wavefront lanes per (block) row that times fastest.  Decisions are cached per pattern signature
(rows, nnz, and the diagonal set), so a re-assembled operator on the same mesh reuses them."""
from __future__ import annotations

import torch

_CACHE: dict = {}

CANDIDATE_LANES = (1, 2, 4, 8, 16, 32)


def _time_spmv(op, reps=20) -> float:
    x = torch.ones(op.n, dtype=torch.float32, device=op.device)
    op.mv(x)
    torch.cuda.synchronize(op.device)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        op.mv(x)
    t1.record()
    torch.cuda.synchronize(op.device)
    return t0.elapsed_time(t1) / reps


def choose_format(A, allow=("dia", "bsr3", "csr")):
    """(format, lanes) for the operator ``A`` (a :class:`jno_amd.fem.CSR`): DIA when it applies
    with a reasonable fill, else the fastest of BSR3 (when it applies) and CSR over the candidate
    lanes per (block) row, timed on the device."""
    from .operator import SolverOperator

    nd, offs = SolverOperator.analyze(A)
    key = (A.shape[0], A.nnz, nd, tuple(offs), tuple(allow))
    if key in _CACHE:
        return _CACHE[key]
    n = A.shape[0]
    choice = None
    if nd > 0 and "dia" in allow:
        # DIA streams 2*ndiag-1 diagonals of n values; accept it unless the fill ratio is poor
        stored = n * (2 * nd - 1)
        if A.nnz >= 0.35 * stored or stored <= 2 * A.nnz:
            choice = ("dia", 1)
    if choice is None:
        best = None
        fmts = []
        if "bsr3" in allow and SolverOperator.bsr3_blocks(A) >= 0:
            fmts.append("bsr3")
        if "csr" in allow:
            fmts.append("csr")
        for fmt in fmts:
            op = SolverOperator(A, fmt=fmt, lanes=1)  # one layout; lanes only change the launch
            for L in CANDIDATE_LANES:
                op.op.lanes = op.lanes = L
                t = _time_spmv(op)
                if best is None or t < best[0]:
                    best = (t, fmt, L)
            del op
        choice = (best[1], best[2])
    _CACHE[key] = choice
    return choice


# Launch shapes of the PCG iteration kernels tried by tune_launch: (block, grid_cap, k1_vector, split)
CANDIDATE_SHAPES = tuple(
    (blk, cap, vec, split)
    for blk in (256, 512, 1024)
    for cap in (256, 512, 768, 1024)
    for vec in (0, 1, 2, 3)
    for split in (0, 1, 3, 4, 5)
    # vec 3 is LDS staging in mode 4 and, in mode 5, the register-resident one-kernel iteration with
    # pipelined tiles (256 and 512 threads); vec 2 is tried only as the register-resident one-kernel
    # variant with late far loads, and cap 768 only with those two
    if (vec != 3 or split == 4 or (split == 5 and blk <= 512))
    and (vec != 2 or (split == 5 and blk <= 512))
    and (cap != 768 or (split == 5 and vec in (2, 3)))
    and (split != 5 or vec in (0, 2, 3))
)
_LAUNCH_CACHE: dict = {}


def tune_launch(A, b, precision: str = "mixed", iters: int = 100, shapes=CANDIDATE_SHAPES):
    """Time ``iters`` bounded PCG iterations on ``(A, b)`` for each launch shape and attach the
    fastest to A's solver operator for ``precision`` (``SolverOperator.tuning``, a per-call
    :class:`jno_amd.solve.Tuning`: later solves on that operator run it; nothing process-wide
    changes).  Returns ``(block, grid_cap, k1_vector, split)``; cached per pattern signature and
    precision, so a re-assembled operator on the same mesh gets the shape without new timings."""
    import time

    from . import precond
    from .solve import LinearSolver, Tuning

    fmt, lanes = choose_format(A)
    key = (A.shape[0], A.nnz, fmt, lanes, precision)
    op = LinearSolver("cg", 1e-10, 0.0, None, precision=precision).operator(A)
    base = op.tuning if getattr(op, "tuning", None) is not None else Tuning.defaults()
    if key not in _LAUNCH_CACHE:
        M = precond.jacobi()
        best = None
        for shape in shapes:
            sv = LinearSolver("cg", 1e-30, 0.0, iters, precision=precision, replace_factor=1e-30, guard=False,
                              tuning=base.with_shape(precision, *shape))
            sv(A, b, M=M)  # warm
            torch.cuda.synchronize(A.device)
            t0 = time.perf_counter()
            sv(A, b, M=M)
            torch.cuda.synchronize(A.device)
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, tuple(shape))
        _LAUNCH_CACHE[key] = best[1]
    op.tuning = base.with_shape(precision, *_LAUNCH_CACHE[key])
    return _LAUNCH_CACHE[key]


BATCH_SHAPES = ((4096, 4096), (4096, 2048), (2048, 4096), (8192, 4096))


def tune_batch(systems, chunks=(8, 16), shapes=BATCH_SHAPES, **solver_kw) -> dict:
    """Launch shape for :class:`jno_amd.batch.BatchSolver` on ``systems`` (a list of ``(A, b)``):
    steps per hipGraph replay from ``chunks`` and the one-kernel step's (rows_per_entry, grid) from
    ``shapes``.  Each candidate is timed on one whole batched solve (after a warm solve that captures
    its graph) and the fastest is returned as BatchSolver keyword arguments.  Longer chunks mean
    fewer host polls but more step launches after the last item converged; longer entries mean
    fewer partial sums for each entry to reduce."""
    import time

    from .batch import BatchSolver

    best = None
    for rpe, grid in shapes:
        for c in chunks:
            kw = {"steps_per_graph": int(c), "rows_per_entry": int(rpe), "grid": int(grid)}
            bs = BatchSolver(**kw, **solver_kw)
            bs(systems)  # warm: graph capture, operator layouts
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bs(systems)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, kw)
    return best[1]
