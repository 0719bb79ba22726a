"""The committed golden fixtures agree with the oracle, and the fixture generator's shortcuts
(node-level pattern + component expansion) agree with the literal compress_plan restatement."""
import json
import os
import sys

import numpy as np
import pytest

from oracle import fem_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden as MG  # noqa: E402


def test_p1_pattern_fixture():
    g = np.load(os.path.join(HERE, "p1_patterns.npz"))
    for n in (2, 4, 16):
        p, t = O.kuhn_box(n)
        rp, col, _, _ = O.pattern_struct(t, 1)
        assert np.array_equal(g[f"rowptr_{n}"], rp) and np.array_equal(g[f"col_{n}"], col)


def test_p2_fixture():
    g = np.load(os.path.join(HERE, "p2v3_pattern_n1.npz"))
    p, t = O.kuhn_box(1)
    p2, c2 = O.promote_p2(p, t)
    assert np.array_equal(g["cells"], c2) and np.array_equal(g["points"], p2)
    rp, col, _, _ = O.pattern_struct(c2, 3)
    assert np.array_equal(g["rowptr"], rp) and np.array_equal(g["col"], col)


def test_c1_solution_fixture():
    g = np.load(os.path.join(HERE, "c1_solution.npz"))
    p, t = O.kuhn_box(16)
    A, b, d = O.poisson_p1_system(p, t, O.k_sin(p, t))
    assert np.array_equal(g["b"], b)
    u = O.direct_solve(A, b)
    assert np.linalg.norm(g["u"] - u) / np.linalg.norm(u) < 1e-13


@pytest.mark.parametrize("n,vec", [(2, 1), (3, 1), (1, 3), (2, 3)])
def test_generator_shortcuts_equal_compress_plan(n, vec):
    p, t = O.kuhn_box(n)
    if vec == 3:
        p, t = O.promote_p2(p, t)
    rp, col, _, _ = O.pattern_struct(t, vec)
    nrp, ncol = MG.node_pattern(t, p.shape[0])
    erp, ecol = MG.expand_vec(nrp, ncol, vec)
    assert np.array_equal(erp, rp) and np.array_equal(ecol, col)


def test_sha_fixture_counts():
    meta = json.load(open(os.path.join(HERE, "pattern_sha256.json")))
    assert meta["p1_n64"]["nnz"] == 4_018_753
    assert meta["p1_n32"]["nnz"] == 513_313
    p, t = O.kuhn_box(4)
    p2, c2 = O.promote_p2(p, t)
    assert meta["p2v3_n4"]["cells_sha256"] == MG.sha(c2)


def test_p_eager_sha_fixture_counts():
    # the committed P_eager hashes carry the closed-form entry counts
    meta = json.load(open(os.path.join(HERE, "pattern_sha256.json")))
    for n in (32, 64):
        assert meta[f"p1_eager_n{n}"]["nse"] == O.p_eager_nse_kuhn(n)
    p, t = O.kuhn_box(32)
    A, _, _ = O.poisson_p1_system(p, t, O.k_sin(p, t))
    assert MG.sha(O.bcoo_view(A)[1]) == meta["p1_eager_n32"]["indices_sha256"]
