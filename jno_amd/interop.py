"""DLPack interop: operators and vectors in and out of the solver slot without copies, for callers
that hold JAX (or CuPy, NumPy) arrays.

The reference's ``linear=`` slot calls ``fn(op, b, *, M=None, x0=None) -> x`` with a
``LinearOperator`` over a BCOO (jno/utils/solver/solver_api.py:243-297, :137-147; a bare callable
drops in, tests/test_fem_solver_slots.py:94-108).  :func:`linear_slot` returns such a callable.
Its operator may be a :class:`jno_amd.fem.CSR`, a BCOO (``data``, ``indices``, ``shape``: JAX's or
:class:`jno_amd.fem.BCOO`), an object with ``bcoo`` = ``(data, idx)`` and ``shape`` (the reference's
LinearOperator), or a ``(data, idx, shape)`` tuple; its vectors may be
any DLPack exporter (``__dlpack__``).  The solution is a torch tensor, itself a DLPack exporter:
``jax.numpy.from_dlpack(x)`` on the JAX side.
"""
from __future__ import annotations

from typing import Optional

import torch

from .fem import CSR


def as_tensor(x, device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """A torch view of ``x``: torch tensors as they are, DLPack exporters without a copy (when the
    device matches), anything else through ``torch.as_tensor``."""
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(x)
    if device is not None and t.device != torch.device(device):
        t = t.to(device)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t


def csr_from_bcoo(data, idx, shape, device: Optional[torch.device] = None) -> CSR:
    """BCOO triplets ``data`` (nse,) and ``idx`` (nse, 2) (row, col), in any order and with
    duplicates (the reference's parametric operator appends the Dirichlet unit diagonal as extra
    (d, d, 1) triplets, jno/utils/solver/fem_utils.py:3332-3359) -> CSR sorted by (row, col) with
    the duplicates summed (segment_sum, jno/utils/solver/fem_native.py:2318-2329)."""
    data = as_tensor(data, device, torch.float64).reshape(-1)
    idx = as_tensor(idx, data.device).reshape(-1, 2).to(torch.int64)
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows != n_cols:
        raise ValueError(f"the solver slot takes square operators, got shape {tuple(shape)}")
    if idx.shape[0] != data.shape[0]:
        raise ValueError(f"{data.shape[0]} values for {idx.shape[0]} index pairs")
    if idx.numel() and (idx.min() < 0 or idx[:, 0].max() >= n_rows or idx[:, 1].max() >= n_cols):
        raise ValueError("BCOO index out of range")
    key = idx[:, 0] * n_cols + idx[:, 1]
    order = torch.sort(key, stable=True).indices
    uniq, inverse = torch.unique_consecutive(key[order], return_inverse=True)
    vals = torch.zeros(uniq.numel(), dtype=torch.float64, device=data.device)
    vals.index_add_(0, inverse, data[order])
    rows = uniq // n_cols
    counts = torch.bincount(rows, minlength=n_rows)
    rowptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=data.device)
    rowptr[1:] = torch.cumsum(counts, 0)
    return CSR(rowptr.to(torch.int32), (uniq % n_cols).to(torch.int32), vals, (n_rows, n_cols))


def to_csr(op, device: Optional[torch.device] = None) -> CSR:
    """A :class:`CSR` from what the ``linear=`` slot may receive (see the module docstring)."""
    if isinstance(op, CSR):
        return op
    if hasattr(op, "indices") and hasattr(op, "data") and hasattr(op, "shape"):
        # a BCOO itself (jax.experimental.sparse.BCOO or jno_amd.fem.BCOO)
        return csr_from_bcoo(op.data, op.indices, op.shape, device)
    if hasattr(op, "bcoo"):
        data, idx = op.bcoo
        return csr_from_bcoo(data, idx, op.shape, device)
    if isinstance(op, (tuple, list)) and len(op) == 3:
        return csr_from_bcoo(*op, device=device)
    raise TypeError("the operator must be a jno_amd CSR, a BCOO (.data, .indices, .shape), an object with "
                    ".bcoo = (data, idx) and .shape, or a (data, idx, shape) tuple")


def linear_slot(solver=None, device: Optional[torch.device] = None):
    """A ``fn(op, b, *, M=None, x0=None) -> x`` callable for the reference's ``linear=`` slot, on
    ``solver`` (default :func:`jno_amd.solve.cg` at tol 1e-10).  Operators and vectors arrive by
    DLPack (or as torch tensors) and are moved to ``device`` (default ``cuda:0``) when they live
    elsewhere; ``M`` is passed to the solver unchanged (None, a Jacobi spec or a diagonal)."""
    from .solve import cg

    solver = solver if solver is not None else cg(tol=1e-10)
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)

    def fn(op, b, *, M=None, x0=None):
        A = to_csr(op, dev)
        if A.device != dev:
            A = CSR(A.rowptr.to(dev), A.col.to(dev), A.vals.to(dev), A.shape)
        bt = as_tensor(b, dev, torch.float64).reshape(-1).contiguous()
        x0t = None if x0 is None else as_tensor(x0, dev, torch.float64).reshape(-1).contiguous()
        return solver(A, bt, M=M, x0=x0t)

    fn.solver = solver
    return fn
