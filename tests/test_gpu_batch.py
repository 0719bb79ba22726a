"""Batched solves: uneven meshes each with its own operator, and many right-hand sides on one
operator, against the oracle's direct solves (<= 1e-9) and the single-system solver."""
import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _systems(ns, seed0=100, fmt_csr=()):
    import jno_amd as J

    out = []
    for i, n in enumerate(ns):
        m = J.kuhn_box(n)
        s = J.fem_poisson(m)
        k = O.k_lognormal(m.n_cells, seed0 + i)
        A, b = s.evaluate({"k": torch.tensor(k, device="cuda")})
        out.append((n, k, A, b))
    return out


@pytest.mark.parametrize("graph,one", [(0, True), (4, True), (0, False), (4, False)])
def test_uneven_batch_matches_oracle(dev, graph, one):
    from jno_amd.batch import BatchSolver

    ns = [3, 5, 8, 12, 4, 9]
    sy = _systems(ns)
    bs = BatchSolver(tol=1e-10, steps_per_graph=graph, one_kernel=one)
    xs = bs([(A, b) for _, _, A, b in sy])
    # Kuhn lattices on their Jacobi-scaled DIA layout take the one-kernel step when asked to
    assert bs.last_one_kernel == one
    for (n, k, A, b), x, st in zip(sy, xs, bs.last_stats):
        p, t = O.kuhn_box(n)
        Ao, bo, d = O.poisson_p1_system(p, t, k)
        xd = O.direct_solve(Ao, bo)
        assert _rel(_np(x), xd) <= 1e-9, (n, st)
        assert st["relres"] <= 1e-10, st
    # larger systems need more iterations; each system kept its own count
    its = [s["iterations"] for s in bs.last_stats]
    assert its[3] > its[0]


@pytest.mark.parametrize("rpe,grid,graph", [(2048, 8, 6), (4096, 0, 8), (8192, 24, 3), (2048, 2048, 0)])
def test_one_kernel_batch_launch_shapes(dev, rpe, grid, graph):
    # entries of several tiles, a grid smaller than the entry count (grid-stride within the XCD
    # slabs), an odd chunk (rounded up to even: launches alternate the state parity), direct
    # launches; systems of 1 to 58 tiles of 2048 rows, the largest (49^3 = 117,649 rows) with
    # offsets 1, 49 and 2401
    from jno_amd.batch import BatchSolver

    ns = [3, 24, 7, 48, 13]
    sy = _systems(ns, seed0=300)
    bs = BatchSolver(tol=1e-10, steps_per_graph=graph, rows_per_entry=rpe, grid=grid)
    xs = bs([(A, b) for _, _, A, b in sy])
    assert bs.last_one_kernel
    for (n, k, A, b), x, st in zip(sy, xs, bs.last_stats):
        assert st["relres"] <= 1e-10, (n, st)
        if n <= 13:
            p, t = O.kuhn_box(n)
            Ao, bo, _ = O.poisson_p1_system(p, t, k)
            assert _rel(_np(x), O.direct_solve(Ao, bo)) <= 1e-9, (n, st)
        else:
            # larger ones against the three-kernel step and the single-system solver
            from jno_amd import precond, solve

            u = solve.cg(tol=1e-10, precision="fp64")(A, b, M=precond.jacobi())
            assert _rel(_np(x), _np(u)) <= 1e-9, (n, st)


def test_one_kernel_batch_rejects_bad_shapes(dev):
    from jno_amd.batch import BatchSolver

    sy = _systems([4])
    for kw in ({"rows_per_entry": 3000}, {"grid": 12}, {"rows_per_entry": 2 * 65536}):
        with pytest.raises(ValueError, match="jno_batch_tuning"):
            BatchSolver(tol=1e-10, **kw)([(sy[0][2], sy[0][3])])


def test_batch_with_csr_items_and_shared_operator(dev):
    from jno_amd.batch import BatchSolver
    from jno_amd.operator import SolverOperator

    sy = _systems([6, 7])
    n, k, A, b = sy[0]
    # force a CSR layout on the second system
    A2 = sy[1][2]
    A2._cache[("solver_op", "auto", True)] = SolverOperator(A2, fmt="csr", lanes=4)
    rng = np.random.default_rng(0)
    rhs = [torch.tensor(rng.standard_normal(A.shape[0]), device="cuda") for _ in range(5)]
    bs = BatchSolver(tol=1e-10, steps_per_graph=8)
    xs = bs([(A, r) for r in rhs] + [(A2, sy[1][3])])
    # a CSR item: the whole batch takes the three-kernel step
    assert not bs.last_one_kernel
    p, t = O.kuhn_box(n)
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    for r, x in zip(rhs, xs[:5]):
        assert _rel(_np(x), O.direct_solve(Ao, _np(r))) <= 1e-9
    p2, t2 = O.kuhn_box(7)
    Ao2, bo2, _ = O.poisson_p1_system(p2, t2, sy[1][1])
    assert _rel(_np(xs[5]), O.direct_solve(Ao2, bo2)) <= 1e-9


def test_batch_zero_rhs_and_cap(dev):
    from jno_amd.batch import BatchSolver

    sy = _systems([5, 6])
    A, b = sy[0][2], sy[0][3]
    bs = BatchSolver(tol=1e-10)
    xs = bs([(A, torch.zeros_like(b)), (sy[1][2], sy[1][3])])
    assert not _np(xs[0]).any() and bs.last_stats[0]["iterations"] == 0
    with pytest.raises(RuntimeError, match="did not solve"):
        BatchSolver(tol=1e-10, maxiter=2)([(A, b)])


def test_c4_batch_256_uneven_and_64_rhs(dev):
    import time

    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.batch import BatchSolver

    rng = np.random.default_rng(4)
    ns = rng.integers(24, 65, size=256)
    systems = []
    for i, n in enumerate(ns):
        m = J.kuhn_box(int(n))
        s = J.fem_poisson(m)
        k = torch.tensor(O.k_lognormal(m.n_cells, 1000 + i), device="cuda")
        systems.append(s.evaluate({"k": k}))
    ndof = sum(A.shape[0] for A, _ in systems)
    bs = BatchSolver(tol=1e-10, steps_per_graph=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs = bs(systems)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert all(s["relres"] <= 1e-10 for s in bs.last_stats)
    assert bs.last_one_kernel
    # spot-check against the single-system mixed solver
    single = solve.cg(tol=1e-10, precision="mixed")
    for i in (0, int(np.argmax(ns)), int(np.argmin(ns))):
        A, b = systems[i]
        assert _rel(_np(xs[i]), _np(single(A, b, M=precond.jacobi()))) <= 1e-9
    its = [s["iterations"] for s in bs.last_stats]
    print(f"\nC4 batch: 256 systems, {ndof} dofs, {dt:.3f} s, iterations {min(its)}..{max(its)}")
    # 64 N(0,1) right-hand sides on one 96^3 operator
    m = J.kuhn_box(96)
    s = J.fem_poisson(m)
    A, b = s.evaluate({"k": torch.tensor(O.k_lognormal(m.n_cells, 7), device="cuda")})
    g = torch.Generator(device="cpu").manual_seed(0)
    rhs = [torch.randn(A.shape[0], generator=g, dtype=torch.float64).to("cuda") for _ in range(64)]
    for r in rhs:
        r[s.dmask.bool()] = 0.0
    t0 = time.perf_counter()
    xs = bs([(A, r) for r in rhs])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert all(st["relres"] <= 1e-10 for st in bs.last_stats)
    # N(0,1) right-hand sides are rough: at rtol 1e-10 a solution carries up to cond(A) * 1e-10 of
    # error, more than 1e-9 here.  The bar is therefore relative to an FP64 solve at rtol 1e-13: the
    # batched solutions must be as accurate as the single-system solver's at the same tolerance.
    tight = solve.cg(tol=1e-13, precision="fp64")
    errs = []
    for j in (0, 63):
        ut = _np(tight(A, rhs[j], M=precond.jacobi()))
        e_batch = _rel(_np(xs[j]), ut)
        e_single = _rel(_np(single(A, rhs[j], M=precond.jacobi())), ut)
        errs.append((e_batch, e_single))
        assert e_batch <= max(1e-9, 3.0 * e_single), (j, e_batch, e_single)
    print(f"C4 64 rhs on 96^3: {dt:.3f} s; error vs rtol-1e-13 FP64 (batched, single) {errs}")


def test_batch_nonfinite_rhs_is_reported(dev):
    # a NaN in one right-hand side makes its true residual NaN: that item stops as non-finite (not
    # as converged) and the call returns JNO_ENONFINITE; the other item still converges
    from jno_amd import _lib
    from jno_amd.batch import BatchSolver

    sy = _systems([5, 6])
    bnan = sy[0][3].clone()
    bnan[7] = float("nan")
    bs = BatchSolver(tol=1e-10)
    with pytest.raises(RuntimeError, match="did not solve 1 of 2"):
        bs([(sy[0][2], bnan), (sy[1][2], sy[1][3])])
    assert bs.last_status == _lib.JNO_ENONFINITE
    assert not np.isfinite(bs.last_stats[0]["relres"])
    assert bs.last_stats[1]["relres"] <= 1e-10


def test_batch_and_single_solver_on_one_p2_operator(dev):
    # the single mixed solver caches a BSR3 layout for a vec-3 operator; the batch (DIA and CSR
    # items only) keeps its own layout, in either call order, and both match the oracle
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.batch import BatchSolver
    from jno_amd.elasticity import ElasticitySystem

    n = 2
    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    E = np.exp(0.3 * np.random.default_rng(4).standard_normal(c2.shape[0]))
    Ao, bo, _ = O.elasticity_p2_system(p2, c2, E)
    uo = O.direct_solve(Ao, bo)
    from jno_amd.autotune import choose_format
    from jno_amd.operator import SolverOperator

    s = ElasticitySystem(J.kuhn_box(n))
    for order in ("single_first", "batch_first"):
        A, b = s.evaluate({"E": torch.tensor(E, device="cuda")})
        sv = solve.cg(tol=1e-10, precision="mixed")
        bs = BatchSolver(tol=1e-10)
        if order == "single_first":
            # the single solver's automatic layout as at C5 size, where BSR3 times fastest
            A._cache[("solver_op", "auto", True)] = SolverOperator(A, fmt="bsr3", lanes=16)
            u = sv(A, b, M=precond.jacobi())
            (x,) = bs([(A, b)])
            assert sv.last_stats["operator"]["format"] == "bsr3", sv.last_stats
        else:
            (x,) = bs([(A, b)])
            u = sv(A, b, M=precond.jacobi())
            # the batch's layout did not leak into the single solver's automatic choice
            assert sv.last_stats["operator"]["format"] == choose_format(A)[0], sv.last_stats
        assert bs._op(A).format in ("dia", "csr")
        assert _rel(_np(u), uo) <= 1e-9 and _rel(_np(x), uo) <= 1e-9, order


def test_batch_autotune_picks_and_keeps_a_launch_shape(dev):
    # the first call on a batch signature times the autotuner's candidates and keeps the fastest;
    # a second solver on the same signature reuses it without timing again
    from jno_amd import autotune, batch
    from jno_amd.batch import BatchSolver

    sy = _systems([5, 9, 7], seed0=400)
    bs = BatchSolver(tol=1e-10, autotune=True)
    xs = bs([(A, b) for _, _, A, b in sy])
    kw = bs.last_launch_shape
    assert (kw["rows_per_entry"], kw["grid"]) in autotune.BATCH_SHAPES and kw["steps_per_graph"] in (8, 16)
    assert bs.tuning.rows_per_entry == kw["rows_per_entry"] and bs.steps_per_graph == kw["steps_per_graph"]
    n_tuned = len(batch._TUNED)
    bs2 = BatchSolver(tol=1e-10, autotune=True)
    xs2 = bs2([(A, b) for _, _, A, b in sy])
    assert len(batch._TUNED) == n_tuned and bs2.last_launch_shape == kw
    for (n, k, A, b), x, x2 in zip(sy, xs, xs2):
        p, t = O.kuhn_box(n)
        Ao, bo, _ = O.poisson_p1_system(p, t, k)
        xd = O.direct_solve(Ao, bo)
        assert _rel(_np(x), xd) <= 1e-9 and _rel(_np(x2), xd) <= 1e-9
