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


# ---------------------------------------------------------------------------------------------
# The traction load, pinned independently of the edge-midpoint rule it is written with: every
# quadratic field v is interpolated exactly by P2, so sum_i b_i v(x_i) = t . int_F v dS must hold
# for the consistent load (surface term, jno/utils/solver/fem_native.py:1882-1960).
# ---------------------------------------------------------------------------------------------

# quadratic fields on the face x = 1 and their closed-form integrals over [0,1]^2 in (y, z)
_FACE_FIELDS = (
    (lambda x, y, z: 1.0 + 0 * x, 1.0),
    (lambda x, y, z: z, 0.5),
    (lambda x, y, z: y * y, 1.0 / 3.0),
    (lambda x, y, z: y * z, 0.25),
    (lambda x, y, z: x * z * z - 2.0 * y + 3.0, 1.0 / 3.0 - 1.0 + 3.0),
)


def _face_load(p2, c2, t, weights=None):
    """Nodal load of the traction t on the face x = 1; ``weights`` (6,) replaces the oracle's
    per-facet split (fractions of |F|) to plant a mistake."""
    tris = O.plane_facets_p2(p2, c2, 0, 1.0)
    contrib = O.facet_traction_p2(p2, tris, t)
    if weights is not None:
        p = p2[tris[:, :3]]
        area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        contrib = (area[:, None] * np.asarray(weights)[None, :])[:, :, None] * np.asarray(t, float)[None, None, :]
    b = np.zeros((p2.shape[0], 3))
    for c in range(3):
        np.add.at(b[:, c], tris.reshape(-1), contrib[:, :, c].reshape(-1))
    return b


def _face_identity_error(p2, b, t):
    worst = 0.0
    x, y, z = p2[:, 0], p2[:, 1], p2[:, 2]
    for fn, exact in _FACE_FIELDS:
        lhs = b.T @ fn(x, y, z)  # (3,): sum_i b_i[c] v(x_i)
        worst = max(worst, float(np.abs(lhs - np.asarray(t) * exact).max()))
    return worst


@pytest.mark.parametrize("n", [1, 2, 3])
def test_traction_load_integrates_quadratics_exactly(n):
    p2, c2 = _mesh(n)
    t = (0.3, -0.7, 1.1)
    assert _face_identity_error(p2, _face_load(p2, c2, t), t) <= 1e-14


def test_traction_identity_detects_a_wrong_split():
    # planted in the oracle's split only: vertex and edge weights swapped, and a lumped |F|/6 per
    # node; both keep the total force, and both fail the quadratic identity
    p2, c2 = _mesh(2)
    t = (0.0, 0.0, -1.0)
    for w in ((1 / 3, 1 / 3, 1 / 3, 0, 0, 0), (1 / 6,) * 6):
        b = _face_load(p2, c2, t, weights=w)
        assert abs(b[:, 2].sum() + 1.0) < 1e-14  # the totals check alone passes
        assert _face_identity_error(p2, b, t) > 1e-3


def test_facet_traction_matches_an_independent_triangle_rule():
    # one skew 6-node triangle in space: the split integrates every quadratic as the collapsed
    # (Duffy) Gauss-Legendre rule does, 5 x 5 points, exact to degree 8 on the triangle
    rng = np.random.default_rng(3)
    V = rng.standard_normal((3, 3))
    mids = np.array([0.5 * (V[1] + V[2]), 0.5 * (V[0] + V[2]), 0.5 * (V[0] + V[1])])  # edges (1,2),(0,2),(0,1)
    pts = np.vstack([V, mids])
    tri = np.arange(6)[None, :]
    t = np.array([0.4, -1.3, 2.0])
    w = O.facet_traction_p2(pts, tri, t)[0]  # (6, 3)
    g, gw = np.polynomial.legendre.leggauss(5)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    area = 0.5 * np.linalg.norm(np.cross(V[1] - V[0], V[2] - V[0]))
    coeffs = rng.standard_normal((6, 10))
    for c in coeffs:
        def v(X):
            x, y, z = X[..., 0], X[..., 1], X[..., 2]
            return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z
                    + c[7] * x * y + c[8] * y * z + c[9] * x * z)

        # Duffy: (s, r) in [0,1]^2 -> barycentric (1-s, s(1-r), s r), Jacobian s
        S, R = np.meshgrid(g, g, indexing="ij")
        W = gw[:, None] * gw[None, :] * S
        X = (1 - S)[..., None] * V[0] + (S * (1 - R))[..., None] * V[1] + (S * R)[..., None] * V[2]
        exact = 2.0 * area * float((W * v(X)).sum())
        assert np.allclose(w.T @ v(pts), t * exact, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_vector_traction_recovers_manufactured_displacement(n):
    # re-expresses tests/test_fem_3d.py:364-382 for the P2 elasticity form: with nu = 0 (lambda = 0)
    # and E = 1, u = (x, 0, 0) has sigma = eps = e_x e_x, so it is clamped at x = 0, traction-free on
    # the lateral faces and carries t = (1, 0, 0) on x = 1; P2 contains it, so the discrete solution
    # equals it when (and only when) the load is the consistent one
    p2, c2 = _mesh(n)
    A, b, _ = O.elasticity_p2_system(p2, c2, np.ones(c2.shape[0]), nu=0.0, traction=(1.0, 0.0, 0.0))
    u = O.direct_solve(A, b).reshape(-1, 3)
    exact = np.stack([p2[:, 0], 0 * p2[:, 0], 0 * p2[:, 0]], 1)
    assert np.abs(u - exact).max() <= 1e-12
