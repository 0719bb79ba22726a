"""Autotuner for the solver-operator layout: DIA when it applies, else CSR with the number of
wavefront lanes per row that times fastest.  Decisions are cached per pattern signature
(rows, nnz, and the diagonal set), so a re-assembled operator on the same mesh reuses them."""
from __future__ import annotations

import torch

_CACHE: dict = {}

CANDIDATE_LANES = (1, 2, 4, 8, 16)


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


def choose_format(A):
    """(format, lanes) for the operator ``A`` (a :class:`jno_amd.fem.CSR`)."""
    from .operator import SolverOperator

    nd, offs = SolverOperator.analyze(A)
    key = (A.shape[0], A.nnz, nd, tuple(offs))
    if key in _CACHE:
        return _CACHE[key]
    n = A.shape[0]
    choice = None
    if nd > 0:
        # DIA streams 2*ndiag-1 diagonals of n values; accept it unless the fill ratio is poor
        stored = n * (2 * nd - 1)
        if A.nnz >= 0.35 * stored or stored <= 2 * A.nnz:
            choice = ("dia", 1)
    if choice is None:
        best = None
        for L in CANDIDATE_LANES:
            op = SolverOperator(A, fmt="csr", lanes=L)
            t = _time_spmv(op)
            if best is None or t < best[0]:
                best = (t, L)
        choice = ("csr", best[1])
    _CACHE[key] = choice
    return choice
