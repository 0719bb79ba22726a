"""Write the golden fixtures of tests/golden/ from the oracle (a restatement of the reference
algorithm, see oracle/fem_oracle.py -- not a run of the reference, which needs Python >= 3.11 and
JAX and cannot be imported here).

    python tests/golden/make_golden.py

Fixtures:
  p1_patterns.npz      CSR rowptr/col of the P1 Kuhn lattice pattern at n = 2, 4, 16
  p2v3_pattern_n1.npz  CSR of the P2 vector (3 comps) pattern at n = 1, plus its P2 mesh
  c1_solution.npz      16^3 P1 Poisson, k = 1 + 0.5 sin sin sin: b, dinv, u (scipy direct solve)
  pattern_sha256.json  SHA-256 of rowptr and col (int32 little endian) and nnz for larger lattices;
                       SHA-256 of the BCOO indices (nse x 2 int32, row-major) of the eliminated P1
                       operator with exact zeros dropped (P_eager) at n = 32, 64
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fem_oracle as O  # noqa: E402


def node_pattern(cells, nn):
    """Node-level sorted unique pairs (same set as pattern_struct with vec=1)."""
    T, m = cells.shape
    rows = np.repeat(cells, m, axis=1).reshape(-1).astype(np.int64)
    cols = np.tile(cells, (1, m)).reshape(-1).astype(np.int64)
    key = np.unique(rows * nn + cols)
    r, c = key // nn, key % nn
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=nn))]).astype(np.int64)
    return rowptr, c.astype(np.int32)


def expand_vec(rowptr, col, vec):
    """Dof pattern of the node pattern for vec components (dof = node*vec + comp)."""
    nn = len(rowptr) - 1
    deg = np.diff(rowptr)
    rp = np.concatenate([[0], np.cumsum(np.repeat(deg * vec, vec))]).astype(np.int64)
    parts = []
    for r in range(nn):
        c = col[rowptr[r]: rowptr[r + 1]].astype(np.int64)
        row = (c[:, None] * vec + np.arange(vec)).reshape(-1)
        parts.extend([row] * vec)
    return rp, np.concatenate(parts).astype(np.int32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def main():
    pats = {}
    for n in (2, 4, 16):
        p, t = O.kuhn_box(n)
        rp, col, _, _ = O.pattern_struct(t, 1)
        pats[f"rowptr_{n}"] = rp.astype(np.int32)
        pats[f"col_{n}"] = col.astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "p1_patterns.npz"), **pats)

    p, t = O.kuhn_box(1)
    p2, c2 = O.promote_p2(p, t)
    rp, col, _, _ = O.pattern_struct(c2, 3)
    np.savez_compressed(os.path.join(HERE, "p2v3_pattern_n1.npz"), rowptr=rp.astype(np.int32), col=col,
                        points=p2, cells=c2.astype(np.int32))

    n = 16
    p, t = O.kuhn_box(n)
    k = O.k_sin(p, t)
    A, b, d = O.poisson_p1_system(p, t, k)
    u = O.direct_solve(A, b)
    np.savez_compressed(os.path.join(HERE, "c1_solution.npz"), b=b, dinv=O.jacobi_inv(A), u=u)

    meta = {"generator": "oracle/fem_oracle.py restatement (reference not importable: Python 3.10, no JAX)"}
    for n in (32, 64):
        p, t = O.kuhn_box(n)
        rp, col = node_pattern(t, p.shape[0])
        meta[f"p1_n{n}"] = {"nnz": int(len(col)), "rowptr_sha256": sha(rp), "col_sha256": sha(col)}
    for n in (4, 8):
        p, t = O.kuhn_box(n)
        p2, c2 = O.promote_p2(p, t)
        rp, col = expand_vec(*node_pattern(c2, p2.shape[0]), 3)
        meta[f"p2v3_n{n}"] = {"nnz": int(len(col)), "rowptr_sha256": sha(rp), "col_sha256": sha(col),
                              "cells_sha256": sha(c2)}
    for n in (32, 64):
        # the BCOO of the assembled, eliminated operator (P_eager: exact zeros dropped), k = k_sin
        p, t = O.kuhn_box(n)
        A, _, _ = O.poisson_p1_system(p, t, O.k_sin(p, t))
        _, idx = O.bcoo_view(A)
        meta[f"p1_eager_n{n}"] = {"nse": int(idx.shape[0]), "indices_sha256": sha(idx)}
    with open(os.path.join(HERE, "pattern_sha256.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
