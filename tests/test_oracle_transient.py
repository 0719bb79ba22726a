"""The oracle's theta-method transient step, pinned by the scheme's amplification factor on a
discrete eigenmode, the steady-state limit, the analytic heat decay and a manufactured forced
solution (the reference's own transient checks, restated on the 3D unit cube), and its adjoint by
central finite differences."""
import numpy as np
import pytest
import scipy.linalg

from oracle import fem_oracle as O

PI = np.pi


def _block(n, k=1.0, f=0.0, g=0.0):
    p, t = O.kuhn_box(n)
    kk = np.full(t.shape[0], k) if np.ndim(k) == 0 else np.asarray(k, float)
    A, c, d = O.poisson_p1_system(p, t, kk, f=f, g=g)
    Mz = O.transient_dirichlet(O.mass_matrix_p1(p, t), d)
    return p, t, A, c, d, Mz


def test_transient_mass_edit_and_step_operators():
    p, t, A, c, d, Mz = _block(3, g=0.25)
    Md = Mz.toarray()
    assert np.all(Md[d, :] == 0.0) and np.all(Md[:, d] == 0.0)
    K, E = O.theta_operators(Mz, A, 0.01, 0.5)
    Ad = A.toarray()
    assert np.allclose(K.toarray(), Md + 0.005 * Ad, rtol=0, atol=1e-17)
    assert np.allclose(E.toarray(), Md - 0.005 * Ad, rtol=0, atol=1e-17)
    # a constrained row of the step: theta dt u_next[d] = -(1-theta) dt u[d] + dt g
    assert np.all(K.toarray()[d, d] == 0.005)
    with pytest.raises(ValueError):
        O.theta_operators(Mz[:, ::-1].tocsr(), A, 0.01, 0.5)


@pytest.mark.parametrize("theta", [1.0, 0.5, 0.7])
def test_amplification_factor_on_a_discrete_eigenmode(theta):
    # on A v = lam M v (free dofs) one theta-step multiplies v by
    # R = (1 - (1-theta) dt lam) / (1 + theta dt lam): pins the signs and weights of both operators
    p, t, A, c, d, Mz = _block(4, k=O.k_lognormal(6 * 64, 3))
    free = np.setdiff1d(np.arange(p.shape[0]), d)
    lam, V = scipy.linalg.eigh(A.toarray()[np.ix_(free, free)], Mz.toarray()[np.ix_(free, free)])
    dt = 0.003
    for j in (0, 5, len(lam) - 1):
        v = np.zeros(p.shape[0])
        v[free] = V[:, j]
        u1 = O.theta_march(Mz, A, np.zeros_like(c), v, dt, theta, 1)[-1]
        R = (1.0 - (1.0 - theta) * dt * lam[j]) / (1.0 + theta * dt * lam[j])
        assert np.abs(u1 - R * v).max() < 1e-11 * np.abs(v).max()


def test_steady_state_limit_and_boundary_values():
    # backward Euler with a huge step from a field that carries g lands on the steady solve A u = c
    p, t, A, c, d, Mz = _block(5, k=2.0, f=1.0, g=0.5)
    u0 = np.zeros(p.shape[0])
    u0[d] = 0.5
    u1 = O.theta_march(Mz, A, c, u0, 1e8, 1.0, 1)[-1]
    us = O.direct_solve(A, c)
    assert np.abs(u1 - us).max() < 1e-7
    assert np.array_equal(u1[d], np.full(len(d), 0.5))
    # trapezoidal steps keep the boundary values of a consistent initial field exactly
    u3 = O.theta_march(Mz, A, c, u0, 0.01, 0.5, 3)[-1]
    assert np.abs(u3[d] - 0.5).max() < 1e-14


def test_heat_decay_of_the_fundamental_mode():
    # tests/test_fem_block_step.py:49-62 on the unit cube: sin(pi x) sin(pi y) sin(pi z) decays as
    # exp(-3 pi^2 t); 10 backward-Euler steps to T = 0.02, bar 5e-2 as in the reference test
    p, t, A, c, d, Mz = _block(12)
    u0 = np.sin(PI * p[:, 0]) * np.sin(PI * p[:, 1]) * np.sin(PI * p[:, 2])
    T = 0.02
    exact = np.exp(-3 * PI**2 * T) * u0
    for theta in (1.0, 0.5):
        u = O.theta_march(Mz, A, c, u0, T / 10, theta, 10)[-1]
        assert np.linalg.norm(u - exact) / np.linalg.norm(exact) < 5e-2


def test_forced_heat_converges_to_the_manufactured_solution():
    # tests/test_fem_time_schemes.py:203-230 on the unit cube: f = ss = sin sin sin, u0 = 0 gives
    # u(T) = ss (1 - exp(-3 pi^2 T)) / (3 pi^2); constant and time-varying (f = ss (1 + 3 pi^2 t)
    # through the forcing increment -> u = ss t) sources, second-order convergence in h
    T, N = 0.3, 60
    errs = {}
    for n in (8, 16):
        p, t = O.kuhn_box(n)
        ss = np.sin(PI * p[:, 0]) * np.sin(PI * p[:, 1]) * np.sin(PI * p[:, 2])
        _, _, A, c, d, Mz = _block(n, f=ss)
        ex = ss * (1 - np.exp(-3 * PI**2 * T)) / (3 * PI**2)
        for theta in (1.0, 0.5):
            u = O.theta_march(Mz, A, c, np.zeros_like(ss), T / N, theta, N)[-1]
            errs[(n, theta)] = np.linalg.norm(u - ex) / np.linalg.norm(ex)
        # time-varying: c carries the t = 0 load, the forcing the increment 3 pi^2 t M ss on free rows
        free = np.ones_like(ss)
        free[d] = 0.0
        inc = free * O.p1_load(p, t, ss, p.shape[0]) * 3 * PI**2
        u = O.theta_march(Mz, A, c, np.zeros_like(ss), T / N, 0.5, N, forcing=lambda tt: tt * inc)[-1]
        errs[(n, "tv")] = np.linalg.norm(u - ss * T) / np.linalg.norm(ss * T)
    for key in (1.0, 0.5, "tv"):
        assert errs[(16, key)] < 2e-2, errs
        assert errs[(8, key)] / errs[(16, key)] > 3.5, errs  # O(h^2)


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_transient_adjoint_matches_central_fd(theta):
    # tests/test_fem_solver_transient.py:303 (transient adjoint vs finite differences), restated:
    # L = w . u_N over a 4-step march with nonzero g and a source
    p, t = O.kuhn_box(3)
    T = t.shape[0]
    k = O.k_lognormal(T, 11)
    w = np.random.default_rng(2).standard_normal(p.shape[0])
    dt, N, g = 0.02, 4, 0.3
    u0 = 0.1 * np.random.default_rng(4).standard_normal(p.shape[0])

    def run(kk):
        _, _, A, c, d, Mz = _block(3, k=kk, f=1.0, g=g)
        return O.theta_march(Mz, A, c, u0, dt, theta, N), A, d, Mz

    us, A, d, Mz = run(k)
    grad, gu0 = O.theta_march_grad_k(p, t, Mz, A, us, dt, theta, d, g, w)
    for e in (0, 7, T // 2, T - 1):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        fd = (w @ run(kp)[0][-1] - w @ run(km)[0][-1]) / (2 * h)
        assert abs(grad[e] - fd) <= 1e-6 * max(abs(fd), 1e-3), (e, grad[e], fd)
    # dL/du0 by linearity in u0
    v = np.random.default_rng(5).standard_normal(p.shape[0])
    _, _, A, c, d, Mz = _block(3, k=k, f=1.0, g=g)
    h = 1e-3
    fd0 = (w @ O.theta_march(Mz, A, c, u0 + h * v, dt, theta, N)[-1]
           - w @ O.theta_march(Mz, A, c, u0 - h * v, dt, theta, N)[-1]) / (2 * h)
    assert abs(gu0 @ v - fd0) < 1e-9 * max(1.0, abs(fd0))


def _tv_block(n, k, profile, f=0.0, g=0.0):
    # the eliminated system at coefficient profile(t) * k, and the transient mass of its Dirichlet set
    p, t = O.kuhn_box(n)
    _, _, d = O.poisson_p1_system(p, t, k, f=f, g=g)
    Mz = O.transient_dirichlet(O.mass_matrix_p1(p, t), d)

    def system_at(tt):
        A, c, _ = O.poisson_p1_system(p, t, profile(tt) * k, f=f, g=g)
        return A, c

    return p, t, d, Mz, system_at


def test_time_dependent_march_with_a_constant_operator_is_the_march():
    p, t, d, Mz, system_at = _tv_block(4, O.k_lognormal(6 * 64, 2), lambda tt: 1.0, f=1.0, g=0.2)
    A, c = system_at(0.0)
    u0 = 0.3 * np.random.default_rng(1).standard_normal(p.shape[0])
    for theta in (1.0, 0.5):
        a = O.theta_march_tv(Mz, system_at, u0, 0.01, theta, 4)
        b = O.theta_march(Mz, A, c, u0, 0.01, theta, 4)
        assert max(np.abs(x - y).max() for x, y in zip(a, b)) < 1e-12


@pytest.mark.parametrize("theta", [1.0, 0.5, 0.7])
def test_time_dependent_amplification_uses_the_operator_at_the_step_end(theta):
    # with k(t) = phi(t) k the free block is phi(t) A_ff, so a generalized eigenvector A_ff v = lam M v
    # is multiplied by prod_n (1 - (1-theta) dt phi(t_{n+1}) lam) / (1 + theta dt phi(t_{n+1}) lam):
    # pins the evaluation time (t_{n+1} for both halves) and the theta weights
    n, dt, N = 4, 0.004, 6
    k = O.k_lognormal(6 * n**3, 3)
    phi = lambda tt: 1.0 + 40.0 * tt  # noqa: E731
    p, t, d, Mz, system_at = _tv_block(n, k, phi)
    A1, _ = system_at(0.0)  # phi(0) = 1: the operator of k itself
    free = np.setdiff1d(np.arange(p.shape[0]), d)
    lam, V = scipy.linalg.eigh(A1.toarray()[np.ix_(free, free)], Mz.toarray()[np.ix_(free, free)])
    for j in (0, 7, len(lam) - 1):
        v = np.zeros(p.shape[0])
        v[free] = V[:, j]
        uN = O.theta_march_tv(Mz, system_at, v, dt, theta, N)[-1]
        R = np.prod([(1 - (1 - theta) * dt * phi((i + 1) * dt) * lam[j]) / (1 + theta * dt * phi((i + 1) * dt) * lam[j])
                     for i in range(N)])
        assert np.abs(uN - R * v).max() < 1e-11 * np.abs(v).max(), j
        # the operator at the step start would give another factor (the check can tell them apart)
        R0 = np.prod([(1 - (1 - theta) * dt * phi(i * dt) * lam[j]) / (1 + theta * dt * phi(i * dt) * lam[j])
                      for i in range(N)])
        assert abs(R - R0) > 1e-4 * abs(R), j


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_time_dependent_adjoint_matches_central_fd(theta):
    n = 3
    p, t = O.kuhn_box(n)
    T = t.shape[0]
    k = O.k_lognormal(T, 12)
    phi = lambda tt: 1.0 + 30.0 * tt + 0.5 * np.sin(200.0 * tt)  # noqa: E731
    w = np.random.default_rng(3).standard_normal(p.shape[0])
    dt, N, g = 0.02, 4, 0.3
    u0 = 0.1 * np.random.default_rng(6).standard_normal(p.shape[0])

    def run(kk):
        _, _, d, Mz, system_at = _tv_block(n, kk, phi, f=1.0, g=g)
        return O.theta_march_tv(Mz, system_at, u0, dt, theta, N), d, Mz, system_at

    us, d, Mz, system_at = run(k)
    grad, gu0 = O.theta_march_tv_grad_k(p, t, Mz, system_at, phi, us, dt, theta, d, g, w)
    for e in (0, 9, T // 2, T - 1):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        fd = (w @ run(kp)[0][-1] - w @ run(km)[0][-1]) / (2 * h)
        assert abs(grad[e] - fd) <= 1e-6 * max(abs(fd), 1e-3), (e, grad[e], fd)
    # a profile left out of the contraction would be caught
    grad1, _ = O.theta_march_tv_grad_k(p, t, Mz, system_at, lambda tt: 1.0, us, dt, theta, d, g, w)
    assert np.abs(grad1 - grad).max() > 1e-3 * np.abs(grad).max()
    v = np.random.default_rng(7).standard_normal(p.shape[0])
    h = 1e-3
    fd0 = (w @ O.theta_march_tv(Mz, system_at, u0 + h * v, dt, theta, N)[-1]
           - w @ O.theta_march_tv(Mz, system_at, u0 - h * v, dt, theta, N)[-1]) / (2 * h)
    assert abs(gu0 @ v - fd0) < 1e-9 * max(1.0, abs(fd0))
