"""The oracle's element values, Dirichlet elimination, PCG and adjoint, pinned by closed forms,
invariants, a direct solver and finite differences.  Each test names the reference test it re-expresses."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_oracle as O


def test_p1_stiffness_row_sums_symmetry_and_values():
    p, t = O.kuhn_box(3)
    k = O.k_lognormal(t.shape[0], 1)
    Ke = O.p1_stiffness(p, t, k)
    assert np.allclose(Ke, np.transpose(Ke, (0, 2, 1)), atol=0)
    assert np.abs(Ke.sum(axis=2)).max() < 1e-12 * np.abs(Ke).max()
    # the reference Kuhn tet (path 0->x->y->z) has the textbook 4x4 Laplacian (h=1, k=1, vol 1/6)
    pp = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1.0]])
    K = O.p1_stiffness(pp, np.array([[0, 1, 2, 3]]), [1.0])[0]
    ref = np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]]) / 6.0
    assert np.allclose(K, ref, atol=1e-15)


def test_p1_linear_field_is_exact():
    # re-expresses tests/test_fem_3d.py:84-97: TET4 recovers a linear field to 1e-9
    p, t = O.kuhn_box(5)
    g = 1.0 + 2.0 * p[:, 0] - 3.0 * p[:, 1] + 0.5 * p[:, 2]
    A, b, d = O.poisson_p1_system(p, t, np.full(t.shape[0], 2.5), f=0.0, g=0.0)
    # nonzero Dirichlet values: rebuild with g on the boundary
    n = p.shape[0]
    rp, col, vals = O.assemble(O.p1_stiffness(p, t, np.full(t.shape[0], 2.5)), t, 1, n)
    A0 = O.csr_matrix(rp, col, vals, n)
    dofs = O.boundary_nodes(p, t)
    A1, b1 = O.dirichlet_eliminate(A0, np.zeros(n), dofs, g[dofs])
    u = O.direct_solve(A1, b1)
    assert np.abs(u - g).max() < 1e-9


def test_mass_integrates_volume_and_is_exact_for_linear():
    p, t = O.kuhn_box(4)
    M = O.mass_matrix_p1(p, t)
    one = np.ones(p.shape[0])
    assert abs(one @ M @ one - 1.0) < 1e-13
    x = p[:, 0]
    assert abs(one @ M @ x - 0.5) < 1e-13  # int x dV
    assert abs(x @ M @ x - 1.0 / 3.0) < 1e-13  # int x^2 dV (degree 2: exact)


def test_load_vector_closed_form():
    p, t = O.kuhn_box(4)
    b = O.p1_load(p, t, 1.0, p.shape[0])
    assert abs(b.sum() - 1.0) < 1e-13
    f = 1.0 + p[:, 0]
    bf = O.p1_load(p, t, f, p.shape[0])
    assert abs(bf.sum() - 1.5) < 1e-13


def test_one_hot_element_support_and_linearity():
    # re-expresses tests/test_topology_optimisation.py:210-237
    p, t = O.kuhn_box(2)
    T, n = t.shape[0], p.shape[0]
    plan = O.pattern_struct(t, 1)
    dense = lambda k: O.csr_matrix(*O.assemble(O.p1_stiffness(p, t, k), t, 1, n, plan), n).toarray()  # noqa: E731
    full = dense(np.ones(T))
    acc = sum(dense(np.eye(T)[j]) for j in range(T))
    assert np.abs(acc - full).max() < 1e-12
    j = 5
    kj = dense(np.eye(T)[j])
    mask = np.zeros(n, bool)
    mask[t[j]] = True
    assert np.abs(kj[~mask][:, ~mask]).max() == 0.0
    assert np.abs(kj[np.ix_(mask, mask)]).max() > 0.0


def test_dirichlet_elimination_identities():
    # re-expresses tests/test_fem_dirichlet_symmetry.py:105-160
    p, t = O.kuhn_box(4)
    n = p.shape[0]
    rp, col, vals = O.assemble(O.p1_stiffness(p, t, O.k_lognormal(t.shape[0], 2)), t, 1, n)
    A0 = O.csr_matrix(rp, col, vals, n)
    dofs = O.boundary_nodes(p, t)
    g = np.sin(p[dofs, 0] + 2 * p[dofs, 1])
    b0 = O.p1_load(p, t, 1.0, n)
    A, b = O.dirichlet_eliminate(A0, b0, dofs, g)
    Ad = A.toarray()
    assert np.array_equal(Ad, Ad.T)
    assert np.array_equal(Ad[dofs], np.eye(n)[dofs])
    assert np.array_equal(b[dofs], g)
    assert A.nnz == A0.nnz  # pattern kept; constrained couplings are stored zeros
    # the eliminated system has the same solution as the row-replacement system
    Ar = A0.toarray()
    Ar[dofs] = np.eye(n)[dofs]
    br = b0.copy()
    br[dofs] = g
    assert np.abs(O.direct_solve(A, b) - np.linalg.solve(Ar, br)).max() < 1e-11


def test_pcg_matches_direct_solve_and_iteration_count():
    # C1: 16^3, k = 1 + 0.5 sin sin sin, rtol 1e-10
    p, t = O.kuhn_box(16)
    A, b, d = O.poisson_p1_system(p, t, O.k_sin(p, t))
    x, it = O.pcg(A, b, O.jacobi_inv(A), tol=1e-10)
    xd = O.direct_solve(A, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert it == 64
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10 * 1.0001


def test_pcg_jax_semantics_edge_cases():
    p, t = O.kuhn_box(6)
    A, b, d = O.poisson_p1_system(p, t, O.k_lognormal(t.shape[0], 4))
    x, it = O.pcg(A, np.zeros_like(b), O.jacobi_inv(A))
    assert it == 0 and not x.any()  # r.r = 0 <= tol^2 b.b: no iteration
    x, it = O.pcg(A, b, O.jacobi_inv(A), maxiter=3)
    assert it == 3
    with pytest.raises(RuntimeError, match="did not solve"):
        O.residual_check(A, b, x, rtol=1e-12)
    # x0 given: already the solution -> zero iterations
    xd = O.direct_solve(A, b)
    _, it = O.pcg(A, b, O.jacobi_inv(A), tol=1e-6, x0=xd)
    assert it == 0


def test_jacobi_zero_diagonal_guard():
    # re-expresses tests/test_fem_linear_solvers.py:80-89
    A = sp.csr_matrix(np.array([[2.0, 0.0], [0.0, 0.0]]))
    assert np.array_equal(O.jacobi_inv(A), [0.5, 1.0])


def _loss_and_grad(n, seed_k=0, seed_t=1):
    p, t = O.kuhn_box(n)
    T = t.shape[0]
    k = O.k_lognormal(T, seed_k)
    At, bt, d = O.poisson_p1_system(p, t, O.k_lognormal(T, seed_t))
    ustar = O.direct_solve(At, bt)
    M = O.mass_matrix_p1(p, t)
    A, b, d = O.poisson_p1_system(p, t, k)
    u = O.direct_solve(A, b)
    rhs = M @ (u - ustar)
    rhs[d] = 0.0
    lam = O.direct_solve(A, rhs)
    grad = O.adjoint_grad_k_p1(p, t, u, lam, d)
    return p, t, k, ustar, M, u, grad


def test_adjoint_gradient_matches_central_fd():
    # re-expresses tests/test_fem_solver_krylov.py:382-397 (FD vs reverse mode)
    p, t, k, ustar, M, u, grad = _loss_and_grad(4)
    rng = np.random.default_rng(7)
    for e in rng.choice(t.shape[0], 8, replace=False):
        h = 1e-6 * k[e]
        Ls = []
        for s in (+1, -1):
            kk = k.copy()
            kk[e] += s * h
            A, b, d = O.poisson_p1_system(p, t, kk)
            # defect correction from u keeps the perturbation above the solver noise
            du = O.direct_solve(A, b - A @ u)
            uu = u + du
            Ls.append(0.5 * (uu - ustar) @ (M @ (uu - ustar)))
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * abs(grad[e]) + 1e-14


def test_adjoint_gradient_sign_on_a_compliance():
    # compliance C = b.u falls when any k grows: dC/dk_e = -u_e^T (K_e/k_e) u_e <= 0
    p, t = O.kuhn_box(3)
    A, b, d = O.poisson_p1_system(p, t, np.ones(t.shape[0]))
    u = O.direct_solve(A, b)
    g = O.adjoint_grad_k_p1(p, t, u, u, d)
    assert np.all(g <= 1e-18)


def test_c_restatement_matches_numpy_oracle():
    from oracle import cpu_fem

    n = 12
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 6)
    mask = np.zeros(p.shape[0], np.uint8)
    mask[O.boundary_nodes(p, t)] = 1
    s = cpu_fem.P1System(p, t, k, 1.0, mask)
    A, b, d = O.poisson_p1_system(p, t, k)
    A.sort_indices()
    assert np.array_equal(s.rowptr, A.indptr) and np.array_equal(s.col, A.indices)
    assert np.abs(s.vals - A.data).max() <= 1e-12 * np.abs(A.data).max()
    assert np.abs(s.b - b).max() <= 1e-14
    x, it, rel, _ = s.pcg(1e-10)
    xo, ito = O.pcg(A, b, O.jacobi_inv(A), tol=1e-10)
    assert abs(it - ito) <= 1 and rel <= 1e-10 * 1.001
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9


@pytest.mark.parametrize("n", [2, 4, 8])
def test_bcoo_view_closed_form_and_brute_force(n):
    # P_eager of the eliminated operator: the closed-form nse of the Kuhn lattice, and the
    # nonzero positions of the dense matrix in row-major order (brute force).  Dyadic spacings only:
    # at n = 3 or 6 the oracle's inverse-Jacobian geometry leaves round-off (~1e-17) where the
    # couplings vanish in exact arithmetic, so its P_eager carries extra entries there.
    p, t = O.kuhn_box(n)
    A, b, d = O.poisson_p1_system(p, t, O.k_lognormal(t.shape[0], n), g=0.5)
    data, idx = O.bcoo_view(A)
    assert idx.dtype == np.int32 and idx.shape == (O.p_eager_nse_kuhn(n), 2)
    if n <= 4:
        D = A.toarray()
        rr, cc = np.nonzero(D)  # row-major order
        assert np.array_equal(idx[:, 0], rr) and np.array_equal(idx[:, 1], cc)
        assert np.array_equal(data, D[rr, cc])
    assert np.all(data != 0.0)
    key = idx[:, 0].astype(np.int64) * (n + 1) ** 3 + idx[:, 1]
    assert np.all(np.diff(key) > 0)
