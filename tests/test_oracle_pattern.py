"""The oracle's structural CSR pattern (compress_plan restatement), pinned by brute force and closed forms."""
import numpy as np
import pytest

from oracle import fem_oracle as O


def brute_pattern(cells, vec, ndof):
    """Dense boolean adjacency of every element dof pair; its row-major nonzeros are sorted (row, col)."""
    Adj = np.zeros((ndof, ndof), dtype=bool)
    cd = O.cell_dofs(cells, vec)
    for row in cd:
        Adj[np.ix_(row, row)] = True
    r, c = np.nonzero(Adj)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=ndof))])
    return rowptr, c


@pytest.mark.parametrize("n", [1, 2, 3])
def test_p1_pattern_matches_brute_force(n):
    p, t = O.kuhn_box(n)
    rp, col, inv, nse = O.pattern_struct(t, 1)
    brp, bcol = brute_pattern(t, 1, p.shape[0])
    assert np.array_equal(rp, brp) and np.array_equal(col, bcol)
    assert nse == len(col)


def test_p2_vec3_pattern_matches_brute_force():
    p, t = O.kuhn_box(1)
    p2, c2 = O.promote_p2(p, t)
    rp, col, inv, nse = O.pattern_struct(c2, 3)
    brp, bcol = brute_pattern(c2, 3, 3 * p2.shape[0])
    assert np.array_equal(rp, brp) and np.array_equal(col, bcol)


def kuhn_p1_nnz(n):
    """N + 2E with E = axis edges + one diagonal per voxel face + one body diagonal per voxel."""
    N = (n + 1) ** 3
    E = 3 * n * (n + 1) ** 2 + 3 * n**2 * (n + 1) + n**3
    return N + 2 * E


@pytest.mark.parametrize("n", [2, 3, 5, 8, 16])
def test_p1_nnz_closed_form(n):
    p, t = O.kuhn_box(n)
    assert O.pattern_struct(t, 1)[3] == kuhn_p1_nnz(n)


def test_config_nnz_values():
    # the counts the benchmark configurations quote (16^3, 64^3, 128^3, 96^3)
    assert kuhn_p1_nnz(16) == 66_961
    assert kuhn_p1_nnz(64) == 4_018_753
    assert kuhn_p1_nnz(128) == 31_802_497
    assert kuhn_p1_nnz(96) == 13_465_441


def test_inverse_map_reassembles_triplets():
    p, t = O.kuhn_box(2)
    rp, col, inv, nse = O.pattern_struct(t, 1)
    cd = O.cell_dofs(t, 1)
    rows = np.repeat(cd, 4, axis=1).reshape(-1)
    cols = np.tile(cd, (1, 4)).reshape(-1)
    # every raw (row, col) triplet lands on its own unique slot
    slot_row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    assert np.array_equal(slot_row[inv], rows)
    assert np.array_equal(col[inv], cols)


def test_eager_pattern_is_the_free_seven_point_stencil():
    # compress_eager on a dyadic Kuhn mesh keeps axis couplings between free nodes + Dirichlet diagonals
    n = 8
    p, t = O.kuhn_box(n)
    A, b, d = O.poisson_p1_system(p, t, O.k_lognormal(t.shape[0], 3))
    Ae = O.compress_eager(A)
    expect = (n - 1) ** 3 + 6 * (n - 1) ** 2 * (n - 2) + (6 * n**2 + 2)
    assert Ae.nnz == expect
    r, c = Ae.nonzero()
    off = np.abs(c - r)
    assert set(np.unique(off)) <= {0, 1, n + 1, (n + 1) ** 2}
