"""The oracle's P2 vector elasticity: exactness on quadratic fields, rigid-body null space, symmetry,
traction load and the compliance gradient."""
import numpy as np
import pytest

from oracle import fem_oracle as O


def _mesh(n):
    p, t = O.kuhn_box(n)
    return O.promote_p2(p, t)


def test_energy_of_quadratic_fields_is_exact():
    # re-expresses tests/test_fem_higher_order.py:133-150 (TET10 is exact on quadratics): the
    # assembled bilinear form equals the analytic integral of lam div u div v + 2 mu eps(u):eps(v)
    nu = 0.3
    lam, mu = O.lame(nu)
    p2, c2 = _mesh(2)
    n = 3 * p2.shape[0]
    K = O.csr_matrix(*O.assemble(O.elasticity_p2(p2, c2, np.ones(c2.shape[0]), nu), c2, 3, n), n)
    x, y, z = p2[:, 0], p2[:, 1], p2[:, 2]
    u = np.stack([x * y, z * z - x, y * z + 2 * x * x], 1).reshape(-1)
    v = np.stack([y * y, x * z, x - y * y + z], 1).reshape(-1)
    # analytic gradients (rows: component, cols: d/dx, d/dy, d/dz)
    def gu(X, Y, Z):
        return np.array([[Y, X, 0 * X], [-1 + 0 * X, 0 * X, 2 * Z], [4 * X, Z, Y]])

    def gv(X, Y, Z):
        return np.array([[0 * X, 2 * Y, 0 * X], [Z, 0 * X, X], [1 + 0 * X, -2 * Y, 1 + 0 * X]])

    # tensor Gauss-Legendre on the unit cube, exact for the degree-2 integrand
    g, w = np.polynomial.legendre.leggauss(4)
    g, w = 0.5 * (g + 1), 0.5 * w
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    Gu, Gv = gu(X, Y, Z), gv(X, Y, Z)
    eu = 0.5 * (Gu + np.swapaxes(Gu, 0, 1))
    ev = 0.5 * (Gv + np.swapaxes(Gv, 0, 1))
    dens = lam * np.trace(Gu) * np.trace(Gv) + 2 * mu * np.einsum("ij...,ij...->...", eu, ev)
    exact = float((W * dens).sum())
    assert abs(u @ (K @ v) - exact) <= 1e-12 * abs(exact)


def test_rigid_body_null_space_and_symmetry():
    p2, c2 = _mesh(2)
    n = 3 * p2.shape[0]
    E = np.exp(0.3 * np.random.default_rng(0).standard_normal(c2.shape[0]))
    K = O.csr_matrix(*O.assemble(O.elasticity_p2(p2, c2, E, 0.3), c2, 3, n), n)
    Kd = K.toarray()
    assert np.abs(Kd - Kd.T).max() <= 1e-14 * np.abs(Kd).max()
    modes = []
    for a in range(3):
        t = np.zeros((p2.shape[0], 3))
        t[:, a] = 1.0
        modes.append(t)
    for a, b in ((0, 1), (1, 2), (0, 2)):
        r = np.zeros((p2.shape[0], 3))
        r[:, a], r[:, b] = -p2[:, b], p2[:, a]
        modes.append(r)
    for mde in modes:
        assert np.abs(K @ mde.reshape(-1)).max() <= 1e-13 * np.abs(Kd).max()
    # exactly six zero eigenvalues
    ev = np.linalg.eigvalsh(Kd)
    assert (np.abs(ev) < 1e-10 * ev.max()).sum() == 6


def test_traction_load_totals_and_clamp():
    p2, c2 = _mesh(3)
    E = np.ones(c2.shape[0])
    A, b, d = O.elasticity_p2_system(p2, c2, E)
    # unit traction (0,0,-1) over the unit face x=1: total force = area x t, all on edge nodes
    assert abs(b.reshape(-1, 3).sum(axis=0)[2] + 1.0) < 1e-14
    assert np.abs(b.reshape(-1, 3)[:, :2]).max() == 0.0
    nodes = np.unique(d // 3)
    assert np.all(p2[nodes, 0] == 0.0) and len(nodes) == (2 * 3 + 1) ** 2
    u = O.direct_solve(A, b)
    x, it = O.pcg(A, b, O.jacobi_inv(A), tol=1e-10)
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-9
    # the tip moves down
    right = np.flatnonzero(p2[:, 0] == 1.0)
    assert u.reshape(-1, 3)[right, 2].mean() < 0


def test_compliance_gradient_fd():
    p2, c2 = _mesh(2)
    T = c2.shape[0]
    E = np.exp(0.3 * np.random.default_rng(1).standard_normal(T))
    A, b, d = O.elasticity_p2_system(p2, c2, E)
    u = O.direct_solve(A, b)
    g = O.compliance_grad_p2(p2, c2, E, 0.3, u)
    assert np.all(g < 0)
    for e in (0, 7, T - 1):
        h = 1e-6 * E[e]
        C = []
        for s in (1, -1):
            EE = E.copy()
            EE[e] += s * h
            A2, b2, _ = O.elasticity_p2_system(p2, c2, EE)
            C.append(b2 @ (u + O.direct_solve(A2, b2 - A2 @ u)))
        fd = (C[0] - C[1]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * abs(g[e])


@pytest.mark.parametrize("n", [1, 2])
def test_p2_pattern_is_node_pattern_times_3x3(n):
    p2, c2 = _mesh(n)
    rp, col, _, _ = O.pattern_struct(c2, 3)
    rp1, col1, _, _ = O.pattern_struct(c2, 1)
    assert np.array_equal(np.diff(rp)[::3], 3 * np.diff(rp1))
    assert np.array_equal(col[rp[0]: rp[1]] // 3, np.repeat(col1[rp1[0]: rp1[1]], 3))
