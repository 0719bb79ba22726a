"""Parity of the HIP theta-method transient march (step operators, fused right-hand side, warm-started
PCG per step, reverse sweep for dL/dk) with the oracle's direct-solve march.

Bars: 1e-12 relative for the step operators' values; 1e-8 relative L2 for marched states (each step
solved to rtol 1e-10 against the oracle's direct solves); 1e-6 relative for gradients."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu

PI = np.pi


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _systems(n, k, f=1.0, g=0.0):
    import jno_amd as J

    m = J.kuhn_box(n)
    p, t = O.kuhn_box(n)
    sysg = J.fem_poisson(m, f=f if np.ndim(f) == 0 else torch.as_tensor(f, device="cuda"), g=g)
    A, c, d = O.poisson_p1_system(p, t, k, f=f, g=g)
    Mz = O.transient_dirichlet(O.mass_matrix_p1(p, t), d)
    return m, sysg, p, t, A, c, d, Mz


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_step_operators_match_oracle(dev, theta):
    from jno_amd.transient import SemidiscreteTimeBlock

    n = 6
    k = O.k_lognormal(6 * n**3, 7)
    m, sysg, p, t, A, c, d, Mz = _systems(n, k, g=0.25)
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": torch.as_tensor(k, device="cuda")})
    dt = 0.0123
    K, E = blk.step_operators(dt, theta)
    ev = E.A.vals
    Ko, Eo = O.theta_operators(Mz, A, dt, theta)
    assert np.abs(_np(K.vals) - Ko.data).max() <= 1e-12 * np.abs(Ko.data).max()
    assert np.abs(_np(ev) - Eo.data).max() <= 1e-12 * np.abs(Eo.data).max()
    assert _rel(_np(K.dinv), O.jacobi_inv(Ko)) < 1e-12
    # constrained rows: theta dt on the diagonal and nothing else
    Kg = sp.csr_matrix((_np(K.vals), Ko.indices, Ko.indptr), shape=Ko.shape)
    assert np.all(Kg.diagonal()[d] == theta * dt)
    off = Kg[d].toarray()
    off[np.arange(len(d)), d] = 0.0
    assert np.all(off == 0.0)
    # cached per (dt, theta)
    assert blk.step_operators(dt, theta)[0] is K
    # the step right-hand side on E's DIA layout and on CSR (every lane count) against the oracle
    from jno_amd.operator import SolverOperator

    u = np.random.default_rng(3).standard_normal(p.shape[0])
    ref = O.theta_rhs(Mz, A, c, u, dt, theta)
    ut = torch.as_tensor(u, device="cuda")
    assert E.describe()["format"] == "dia"
    assert _rel(_np(blk.rhs(ut, 0.0, dt, theta)), ref) < 1e-13
    for lanes in (1, 2, 4, 8, 16, 32, 64):
        blk._ops[(dt, theta)] = (K, SolverOperator(E.A, fmt="csr", lanes=lanes, fp32=False))
        assert _rel(_np(blk.rhs(ut, 0.0, dt, theta)), ref) < 1e-13, lanes
    blk._ops[(dt, theta)] = (K, E)


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_march_matches_oracle(dev, theta):
    import jno_amd as J
    from jno_amd.transient import SemidiscreteTimeBlock

    n = 10
    k = O.k_lognormal(6 * n**3, 5)
    m, sysg, p, t, A, c, d, Mz = _systems(n, k, f=1.0, g=0.3)
    u0 = 0.2 * np.random.default_rng(1).standard_normal(p.shape[0])
    dt, N = 0.01, 5
    ref = O.theta_march(Mz, A, c, u0, dt, theta, N)
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": torch.as_tensor(k, device="cuda")}, dt=dt, theta=theta)
    us = blk.march(torch.as_tensor(u0, device="cuda"), N, keep="all")
    for i in (1, N):
        assert _rel(_np(us[i]), ref[i]) < 1e-8, i
    assert all(it > 0 for it in blk.last_iterations)
    # the step operator runs on the index-free diagonal layout; the steady operator of the same
    # plan keeps its own (fewer) diagonals afterwards
    K, _ = blk.step_operators(dt, theta)
    sol = J.solve.cg(tol=1e-10)
    assert sol.operator(K).describe()["format"] == "dia"
    A2, _ = sysg.evaluate({"k": torch.as_tensor(k, device="cuda")})
    nd_k = sol.operator(K).ndiag
    assert sol.operator(A2).ndiag < nd_k


def test_time_varying_forcing_matches_oracle(dev):
    from jno_amd.transient import SemidiscreteTimeBlock

    n = 8
    p, t = O.kuhn_box(n)
    ss = np.sin(PI * p[:, 0]) * np.sin(PI * p[:, 1]) * np.sin(PI * p[:, 2])
    k = np.ones(t.shape[0])
    m, sysg, p, t, A, c, d, Mz = _systems(n, k, f=ss)
    free = np.ones_like(ss)
    free[d] = 0.0
    inc = free * O.p1_load(p, t, ss, p.shape[0]) * 3 * PI**2
    inc_d = torch.as_tensor(inc, device="cuda")
    dt, N = 0.3 / 20, 20
    ref = O.theta_march(Mz, A, c, np.zeros_like(ss), dt, 0.5, N, forcing=lambda tt: tt * inc)
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": torch.ones(t.shape[0], device="cuda", dtype=torch.float64)},
                                            dt=dt, theta=0.5, forcing_vector_fn=lambda tt, args: tt * inc_d)
    u = blk.march(torch.zeros(p.shape[0], device="cuda", dtype=torch.float64), N)
    assert _rel(_np(u), ref[-1]) < 1e-8
    # and the manufactured answer u = ss t within the mesh's discretisation error
    assert _rel(_np(u), ss * 0.3) < 8e-2


@pytest.mark.parametrize("dims", [(10, 10, 10), (300, 3, 3), (255, 3, 3)])
def test_mixed_march_on_eight_diagonal_layouts(dev, dims):
    # the step operator carries the mass diagonals: 8 stored diagonals, with 3 near ones and the
    # far ones inside one window (10^3), 1 near one (nx = 300), or 2 near ones (nx = 255, not an
    # LDS-staged set: the generic kernels run).  (With 2 cells across, every coupling off the x
    # lines touches a constrained node and the nonzero pattern is tridiagonal.)
    import jno_amd as J
    from jno_amd.transient import SemidiscreteTimeBlock

    m = J.kuhn_box(*dims)
    p, t = O.kuhn_box(*dims)
    k = O.k_lognormal(t.shape[0], 9)
    sysg = J.fem_poisson(m, f=1.0, g=0.2)
    A, c, d = O.poisson_p1_system(p, t, k, f=1.0, g=0.2)
    Mz = O.transient_dirichlet(O.mass_matrix_p1(p, t), d)
    u0 = np.full(p.shape[0], 0.2)
    dt, N = 2e-3, 4
    ref = O.theta_march(Mz, A, c, u0, dt, 0.5, N)
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": torch.as_tensor(k, device="cuda")}, dt=dt, theta=0.5)
    sol = J.solve.cg(tol=1e-10, precision="mixed", replace_factor=1e-3)
    u = blk.march(torch.as_tensor(u0, device="cuda"), N, linear_solve=sol)
    assert _rel(_np(u), ref[-1]) < 1e-8
    assert sol.last_stats["operator"]["ndiag"] == 8


def test_heat_decay_of_the_fundamental_mode(dev):
    # tests/test_fem_block_step.py:49-62 on the unit cube (bar 5e-2 as in the reference test)
    import jno_amd as J

    n = 12
    m = J.kuhn_box(n)
    sysg = J.fem_poisson(m, f=0.0)
    p = _np(m.points)
    u0 = np.sin(PI * p[:, 0]) * np.sin(PI * p[:, 1]) * np.sin(PI * p[:, 2])
    T = 0.02
    blk = J.SemidiscreteTimeBlock.from_system(sysg, {"k": 1.0}, dt=T / 10)
    u = blk.march(torch.as_tensor(u0, device="cuda"), 10)
    exact = np.exp(-3 * PI**2 * T) * u0
    assert _rel(_np(u), exact) < 5e-2


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_march_gradient_matches_oracle_adjoint(dev, theta):
    import jno_amd as J

    n = 4
    p, t = O.kuhn_box(n)
    T = t.shape[0]
    k = O.k_lognormal(T, 13)
    g, dt, N = 0.3, 0.02, 4
    m, sysg, p, t, A, c, d, Mz = _systems(n, k, f=1.0, g=g)
    u0 = 0.1 * np.random.default_rng(4).standard_normal(p.shape[0])
    w = np.random.default_rng(2).standard_normal(p.shape[0])
    us = O.theta_march(Mz, A, c, u0, dt, theta, N)
    gk_ref, gu_ref = O.theta_march_grad_k(p, t, Mz, A, us, dt, theta, d, g, w)

    kt = torch.as_tensor(k, device="cuda").requires_grad_(True)
    u0t = torch.as_tensor(u0, device="cuda").requires_grad_(True)
    out = J.theta_march(sysg, kt, u0t, dt, N, theta=theta, solver=J.solve.cg(tol=1e-12))
    assert _rel(_np(out[-1]), us[-1]) < 1e-9
    (out[-1] @ torch.as_tensor(w, device="cuda")).backward()
    assert _rel(_np(kt.grad), gk_ref) < 1e-6
    assert _rel(_np(u0t.grad), gu_ref) < 1e-6


def test_march_gradient_through_saved_states_matches_fd(dev):
    # a loss over several saved states (the trajectory readout of the reference's
    # SemidiscreteTimeBlock.solve(save_ts=...)), against GPU central differences
    import jno_amd as J

    n = 4
    m = J.kuhn_box(n)
    sysg = J.fem_poisson(m, f=1.0, g=0.1)
    T = m.n_cells
    k0 = torch.as_tensor(O.k_lognormal(T, 21), device="cuda")
    u0 = torch.zeros(m.n_nodes, device="cuda", dtype=torch.float64)
    sol = J.solve.cg(tol=1e-13)
    wts = torch.as_tensor(np.random.default_rng(8).standard_normal((3, m.n_nodes)), device="cuda")

    def loss(kk):
        ys = J.theta_march(sysg, kk, u0, 0.05, 6, theta=0.5, save=[2, 4, 6], solver=sol)
        return (ys * wts).sum()

    kt = k0.clone().requires_grad_(True)
    loss(kt).backward()
    for e in (0, T // 3, T - 1):
        h = 1e-6 * float(k0[e])
        kp, km = k0.clone(), k0.clone()
        kp[e] += h
        km[e] -= h
        fd = (float(loss(kp)) - float(loss(km))) / (2 * h)
        assert abs(float(kt.grad[e]) - fd) <= 1e-5 * max(abs(fd), 1e-4), (e, float(kt.grad[e]), fd)


def _tv_oracle(n, k, phi, f=1.0, g=0.0):
    p, t = O.kuhn_box(n)
    _, _, d = O.poisson_p1_system(p, t, k, f=f, g=g)
    Mz = O.transient_dirichlet(O.mass_matrix_p1(p, t), d)

    def system_at(tt):
        A, c, _ = O.poisson_p1_system(p, t, phi(tt) * k, f=f, g=g)
        return A, c

    return p, t, d, Mz, system_at


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_time_dependent_operator_matches_oracle(dev, theta):
    # operator_fn(t, args): the coefficient phi(t) k re-assembled at every step's end time, the
    # lift with it; states against the oracle's direct-solve march of the same time-dependent block
    import jno_amd as J
    from jno_amd.transient import SemidiscreteTimeBlock

    n = 8
    k = O.k_lognormal(6 * n**3, 9)
    phi = lambda tt: 1.0 + 40.0 * tt  # noqa: E731
    g, dt, N = 0.3, 0.01, 5
    p, t, d, Mz, system_at = _tv_oracle(n, k, phi, f=1.0, g=g)
    u0 = 0.2 * np.random.default_rng(2).standard_normal(p.shape[0])
    ref = O.theta_march_tv(Mz, system_at, u0, dt, theta, N)
    m = J.kuhn_box(n)
    sysg = J.fem_poisson(m, f=1.0, g=g)
    kd = torch.as_tensor(k, device="cuda")
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": kd}, coefficient_fn=lambda tt, a: phi(tt) * kd, dt=dt,
                                            theta=theta)
    us = blk.march(torch.as_tensor(u0, device="cuda"), N, keep="all")
    for i in (1, 3, N):
        assert _rel(_np(us[i]), ref[i]) < 1e-8, i
    # the constant-operator march differs (the check sees the time dependence)
    const = SemidiscreteTimeBlock.from_system(sysg, {"k": kd}, dt=dt, theta=theta).march(
        torch.as_tensor(u0, device="cuda"), N)
    assert _rel(_np(const), ref[N]) > 1e-3
    # an operator on another pattern is refused
    bad = SemidiscreteTimeBlock(blk.M, blk.A, blk.c, dmask=blk.dmask, g=blk.g, dt=dt,
                                operator_fn=lambda tt, a: J.fem_poisson(J.kuhn_box(n - 1)).evaluate({"k": 1.0}))
    with pytest.raises(ValueError, match="pattern"):
        bad.step(us[0], 0.0)


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_time_dependent_step_operators_on_the_diagonal_layout(dev, theta):
    # the per-step K, E formed on the DIA layouts of the mass and of A(t) (no CSR pass) equal the
    # CSR path bitwise: values, Jacobi diagonal, the DIA layouts a fill of the CSR values gives, and
    # the step's right-hand side
    import jno_amd as J
    from jno_amd.operator import SolverOperator
    from jno_amd.transient import SemidiscreteTimeBlock

    n, dt = 7, 0.013
    m = J.kuhn_box(n)
    sysg = J.fem_poisson(m, f=1.0, g=0.2)
    kd = torch.as_tensor(O.k_lognormal(m.n_cells, 31), device="cuda")
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": kd}, coefficient_fn=lambda tt, a: (1.0 + 5.0 * tt) * kd,
                                            dt=dt, theta=theta)
    for t_next in (dt, 0.4):
        A, c = blk.operator_at(t_next)
        K, E = blk._form_step(A, dt, theta)
        kop = K._cache[("solver_op", "auto", True)]
        assert kop.format == "dia" and kop.ndiag == 8 and E.format == "dia"
        Kc, Ec = blk._form_step_csr(A, dt, theta)
        assert torch.equal(K.dinv, Kc.dinv)
        assert torch.equal(K.vals, Kc.vals) and torch.equal(E.A.vals, Ec.A.vals)
        ref = SolverOperator(Kc, fmt="dia", fp32=True)
        assert ref.offsets == kop.offsets
        assert torch.equal(kop._keep[0], ref._keep[0]) and torch.equal(kop._keep[1], ref._keep[1])
        eref = SolverOperator(Ec.A, fmt="dia", fp32=False)
        assert eref.offsets == E.offsets and torch.equal(E._keep[0], eref._keep[0])
        u = torch.as_tensor(np.random.default_rng(3).standard_normal(m.n_nodes), device="cuda")
        r1 = blk.rhs(u, t_next - dt, dt, theta, ops=(K, E, c))
        r2 = blk.rhs(u, t_next - dt, dt, theta, ops=(Kc, Ec, c))
        assert torch.equal(r1, r2)


def test_time_dependent_amplification_closed_form(dev):
    # a generalized eigenvector of the free block is multiplied by
    # prod_n (1 - (1-theta) dt phi(t_{n+1}) lam) / (1 + theta dt phi(t_{n+1}) lam) over the march
    import jno_amd as J
    import scipy.linalg
    from jno_amd.transient import SemidiscreteTimeBlock

    n, dt, N, theta = 5, 0.004, 6, 0.5
    k = O.k_lognormal(6 * n**3, 4)
    phi = lambda tt: 1.0 + 40.0 * tt  # noqa: E731
    p, t, d, Mz, system_at = _tv_oracle(n, k, phi, f=0.0)
    A1, _ = system_at(0.0)
    free = np.setdiff1d(np.arange(p.shape[0]), d)
    lam, V = scipy.linalg.eigh(A1.toarray()[np.ix_(free, free)], Mz.toarray()[np.ix_(free, free)])
    sysg = J.fem_poisson(J.kuhn_box(n), f=0.0)
    kd = torch.as_tensor(k, device="cuda")
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": kd}, coefficient_fn=lambda tt, a: phi(tt) * kd, dt=dt,
                                            theta=theta)
    for j in (0, 11):
        v = np.zeros(p.shape[0])
        v[free] = V[:, j]
        uN = blk.march(torch.as_tensor(v, device="cuda"), N, linear_solve=J.solve.cg(tol=1e-12))
        R = np.prod([(1 - (1 - theta) * dt * phi((i + 1) * dt) * lam[j]) / (1 + theta * dt * phi((i + 1) * dt) * lam[j])
                     for i in range(N)])
        # each step solved to 1e-12 of its own right-hand side: errors relative to the initial mode
        assert np.linalg.norm(_np(uN) - R * v) < 1e-10 * np.linalg.norm(v), j


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_time_dependent_march_gradient_matches_oracle_adjoint(dev, theta):
    import jno_amd as J

    n = 4
    p, t = O.kuhn_box(n)
    T = t.shape[0]
    k = O.k_lognormal(T, 17)
    phi = lambda tt: 1.0 + 30.0 * tt + 0.5 * np.sin(200.0 * tt)  # noqa: E731
    g, dt, N = 0.3, 0.02, 4
    _, _, d, Mz, system_at = _tv_oracle(n, k, phi, f=1.0, g=g)
    u0 = 0.1 * np.random.default_rng(5).standard_normal(p.shape[0])
    w = np.random.default_rng(6).standard_normal(p.shape[0])
    us = O.theta_march_tv(Mz, system_at, u0, dt, theta, N)
    gk_ref, gu_ref = O.theta_march_tv_grad_k(p, t, Mz, system_at, phi, us, dt, theta, d, g, w)
    sysg = J.fem_poisson(J.kuhn_box(n), f=1.0, g=g)
    kt = torch.as_tensor(k, device="cuda").requires_grad_(True)
    u0t = torch.as_tensor(u0, device="cuda").requires_grad_(True)
    out = J.theta_march(sysg, kt, u0t, dt, N, theta=theta, solver=J.solve.cg(tol=1e-12), profile=phi)
    assert _rel(_np(out[-1]), us[-1]) < 1e-9
    (out[-1] @ torch.as_tensor(w, device="cuda")).backward()
    assert _rel(_np(kt.grad), gk_ref) < 1e-6
    assert _rel(_np(u0t.grad), gu_ref) < 1e-6


def test_large_march_properties(dev):
    # 64^3 cells (274,625 dofs): energy decay of the homogeneous backward-Euler march, per-step
    # residual of the step solve, Dirichlet rows kept at g, mixed-precision steps agree with FP64
    import jno_amd as J

    n = 64
    m = J.kuhn_box(n)
    sysg = J.fem_poisson(m, f=0.0)
    blk = J.SemidiscreteTimeBlock.from_system(sysg, {"k": 1.0}, dt=1e-3)
    p = _np(m.points)
    u0 = torch.as_tensor(np.sin(PI * p[:, 0]) * np.sin(PI * p[:, 1]) * np.sin(PI * p[:, 2]) + 0.1 * np.cos(7 * p[:, 0]),
                         device="cuda")
    u0[sysg.dmask.bool()] = 0.0
    us = blk.march(u0, 3, keep="all", linear_solve=J.solve.cg(tol=1e-10))
    K, _ = blk.step_operators(1e-3, 1.0)
    Mz_u = [blk.apply_explicit(u, 1e-3, 1.0) for u in us]  # theta = 1: E = Mz
    energy = [float(u @ mu) for u, mu in zip(us, Mz_u)]
    assert all(energy[i + 1] < energy[i] for i in range(3)), energy
    for i in range(3):
        b = blk.rhs(us[i], 0.0, 1e-3, 1.0)
        r = K.mv(us[i + 1]) - b
        assert float(r.norm() / b.norm()) <= 2e-10
        assert float(us[i + 1][sysg.dmask.bool()].abs().max()) == 0.0
    um = blk.march(u0, 3, linear_solve=J.solve.cg(tol=1e-10, precision="mixed"))
    assert float((um - us[-1]).norm() / us[-1].norm()) < 1e-8


def test_transient_argument_errors(dev):
    import jno_amd as J
    from jno_amd.transient import SemidiscreteTimeBlock

    m = J.kuhn_box(3)
    sysg = J.fem_poisson(m)
    blk = SemidiscreteTimeBlock.from_system(sysg, {"k": 1.0})
    with pytest.raises(ValueError, match="theta"):
        blk.step_operators(0.01, 0.0)
    with pytest.raises(ValueError, match="dt > 0"):
        blk.step_operators(-0.01, 1.0)
    with pytest.raises(ValueError):
        blk.step(torch.zeros(m.n_nodes, device="cuda"), 0.0)  # no dt
    with pytest.raises(ValueError):
        blk.step(torch.zeros(m.n_nodes + 1, device="cuda"), 0.0, 0.01)
    other = J.fem_poisson(J.kuhn_box(2)).mass()
    with pytest.raises(ValueError):
        SemidiscreteTimeBlock(other, blk.A, blk.c)
    with pytest.raises(ValueError):
        J.theta_march(sysg, 1.0, torch.zeros(m.n_nodes, device="cuda"), 0.01, 3, save=[4])
