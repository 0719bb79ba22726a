"""Parity of the HIP P2 elasticity path with the oracle: promotion (bit-exact), pattern (bit-exact),
coloring validity, MFMA element matrices + colored scatter (1e-12), traction load, solves (1e-9) and
the compliance gradient (oracle and central FD)."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 2, 3, 8])
def test_promotion_bit_exact(dev, n):
    import jno_amd as J
    from jno_amd.elasticity import promote_p2

    m2 = promote_p2(J.kuhn_box(n))
    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    assert np.array_equal(_np(m2.cells), c2)
    assert np.array_equal(_np(m2.points), p2)


def test_promotion_non_unit_box(dev):
    import jno_amd as J
    from jno_amd.elasticity import promote_p2

    lo, hi = (-0.3, 1.0, 2.0), (0.7, 1.9, 2.5)
    m2 = promote_p2(J.Shape.box(*lo, *hi).structured((2, 3, 4)).domain())
    p2, c2 = O.promote_p2(*O.kuhn_box(2, 3, 4, lo, hi))
    assert np.array_equal(_np(m2.cells), c2) and np.array_equal(_np(m2.points), p2)


def test_p2_boundary_and_pattern(dev):
    import jno_amd as J
    from jno_amd.elasticity import promote_p2

    n = 2
    m2 = promote_p2(J.kuhn_box(n))
    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    assert np.array_equal(np.flatnonzero(_np(m2.boundary_mask())), O.boundary_nodes(p2, c2))
    assert np.array_equal(np.flatnonzero(_np(m2.boundary_mask("left"))), O.boundary_nodes(p2, c2, "left"))
    plan = m2.plan(3)
    rp, col, _, _ = O.pattern_struct(c2, 3)
    assert np.array_equal(_np(plan.rowptr), rp) and np.array_equal(_np(plan.col), col)
    g = np.load(os.path.join(GOLDEN, "p2v3_pattern_n1.npz"))
    plan1 = promote_p2(J.kuhn_box(1)).plan(3)
    assert np.array_equal(_np(plan1.rowptr), g["rowptr"]) and np.array_equal(_np(plan1.col), g["col"])


@pytest.mark.parametrize("n", [4, 8])
def test_p2_pattern_sha256(dev, n):
    import hashlib

    import jno_amd as J
    from jno_amd.elasticity import promote_p2

    meta = json.load(open(os.path.join(GOLDEN, "pattern_sha256.json")))[f"p2v3_n{n}"]
    m2 = promote_p2(J.kuhn_box(n))
    sha = lambda t: hashlib.sha256(np.ascontiguousarray(_np(t), dtype="<i4").tobytes()).hexdigest()  # noqa: E731
    assert sha(m2.cells) == meta["cells_sha256"]
    plan = m2.plan(3)
    assert plan.nnz == meta["nnz"]
    assert sha(plan.rowptr) == meta["rowptr_sha256"] and sha(plan.col) == meta["col_sha256"]


def test_coloring_is_conflict_free(dev):
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem

    s = ElasticitySystem(J.kuhn_box(4))
    cells = _np(s.mesh.cells)
    order = _np(s.coloring.order)
    seen = np.zeros(cells.shape[0], int)
    for a, b in s.coloring.groups():
        grp = order[a:b]
        seen[grp] += 1
        nodes = cells[grp].reshape(-1)
        assert len(np.unique(nodes)) == len(nodes), "two cells of one color share a node"
    assert np.all(seen == 1)
    assert s.coloring.ncolors <= 128


@pytest.mark.parametrize("n", [2, 3])
def test_elasticity_assembly(dev, n):
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem

    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    E = np.exp(0.3 * np.random.default_rng(0).standard_normal(c2.shape[0]))
    s = ElasticitySystem(J.kuhn_box(n))
    A, b = s.evaluate({"E": torch.tensor(E, device="cuda")})
    Ao, bo, d = O.elasticity_p2_system(p2, c2, E)
    Ao.sort_indices()
    assert np.array_equal(_np(A.col), Ao.indices)
    assert np.abs(_np(A.vals) - Ao.data).max() <= 1e-12 * np.abs(Ao.data).max()
    assert np.abs(_np(b) - bo).max() <= 1e-15
    D = _np(A.dense())
    assert np.abs(D - D.T).max() <= 1e-13 * np.abs(D).max()
    # the element-matrix path (MFMA per cell, colored scatter), with and without the per-plan pair
    # table: the same values as each other bitwise, and as the class-table row pull to rounding
    assert s.pair_rank is not None
    assert s._class_geometry() is not None  # classes by bitwise shape (fragmented at spacing 1/3)
    s.class_assembly = False
    A1, _ = s.evaluate({"E": torch.tensor(E, device="cuda")})
    table = s.pair_rank
    s.pair_rank = None
    A2, _ = s.evaluate({"E": torch.tensor(E, device="cuda")})
    s.pair_rank, s.class_assembly = table, True
    assert torch.equal(A1.vals, A2.vals)
    assert np.abs(_np(A.vals) - _np(A1.vals)).max() <= 1e-14 * np.abs(_np(A1.vals)).max()


def test_bsr3_layout(dev):
    """The 3x3-block layout reproduces the CSR product in FP64 and FP32 for every lane count,
    and is refused for a scalar operator."""
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem
    from jno_amd.operator import SolverOperator

    n = 3
    E = np.exp(0.3 * np.random.default_rng(2).standard_normal(6 * n**3))
    s = ElasticitySystem(J.kuhn_box(n))
    A, b = s.evaluate({"E": torch.tensor(E, device="cuda")})
    assert SolverOperator.bsr3_blocks(A) * 9 == A.nnz
    x = torch.randn(A.shape[0], dtype=torch.float64, device="cuda")
    ref = _np(A.mv(x))
    for L in (1, 2, 4, 8, 16, 32, 64):
        op = SolverOperator(A, fmt="bsr3", lanes=L)
        assert _rel(_np(op.mv(x)), ref) < 1e-14
        assert _rel(_np(op.mv(x.float())).astype(np.float64), ref) < 1e-6
    # a scalar P1 operator with 3k rows is not block-structured
    m1 = J.kuhn_box(2)
    A1, _ = J.fem_poisson(m1).evaluate({"k": torch.ones(m1.n_cells, dtype=torch.float64, device="cuda")})
    assert A1.shape[0] % 3 == 0 and SolverOperator.bsr3_blocks(A1) == -1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bsr3_solves_every_iteration_mode(dev, precision):
    """PCG on the BSR3 layout (two-kernel mode falls back to the split iteration) matches the
    oracle's direct solve."""
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.elasticity import ElasticitySystem

    n = 3
    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    E = np.exp(0.3 * np.random.default_rng(3).standard_normal(c2.shape[0]))
    s = ElasticitySystem(J.kuhn_box(n))
    A, b = s.evaluate({"E": torch.tensor(E, device="cuda")})
    Ao, bo, d = O.elasticity_p2_system(p2, c2, E)
    uo = O.direct_solve(Ao, bo)
    base = solve.Tuning.defaults()
    for mode in (0, 1, 2, 3, 4, 5):
        for blk, cap in ((256, 8), (1024, 512)):
            sv = solve.cg(tol=1e-10, precision=precision, format="bsr3",
                          tuning=base.with_shape(precision, blk, cap, 0, mode))
            u = sv(A, b, M=precond.jacobi())
            st = sv.last_stats
            assert st["operator"]["format"] == "bsr3"
            assert st["relres"] <= (1e-10 if precision == "mixed" else 2e-10), (mode, blk, cap, st)
            assert _rel(_np(u), uo) <= 1e-9, (mode, blk, cap, st)


@pytest.mark.parametrize("operator", ["ebe", "assembled"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_elasticity_solve_and_compliance_gradient(dev, precision, operator):
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.elasticity import ElasticitySystem, compliance

    n = 4
    p2, c2 = O.promote_p2(*O.kuhn_box(n))
    T = c2.shape[0]
    E = np.exp(0.3 * np.random.default_rng(1).standard_normal(T))
    s = ElasticitySystem(J.kuhn_box(n))
    sv = solve.cg(tol=1e-10, precision=precision)
    Et = torch.tensor(E, device="cuda", requires_grad=True)
    Cval, u = compliance(s, Et, sv, operator=operator)
    Cval.backward()
    assert (sv.last_stats["operator"]["format"] == "ebe") == (operator == "ebe"), sv.last_stats
    Ao, bo, d = O.elasticity_p2_system(p2, c2, E)
    uo = O.direct_solve(Ao, bo)
    assert _rel(_np(u), uo) <= 1e-9, sv.last_stats
    assert abs(float(Cval.detach()) - bo @ uo) <= 1e-9 * abs(bo @ uo)
    go = O.compliance_grad_p2(p2, c2, E, 0.3, uo)
    assert _rel(_np(Et.grad), go) <= 1e-8
    # central FD (h = 1e-6 E_e) with defect-correction solves to 1e-12 of the defect, against the
    # adjoint gradient of a tight solve (so the rtol-1e-10 solver error is not counted as FD error)
    fine = solve.cg(tol=1e-12)
    A0, b0 = s.evaluate({"E": torch.tensor(E, device="cuda")})
    u0 = fine(A0, b0, M=precond.jacobi())
    u0 = u0 + fine(A0, b0 - A0.mv(u0), M=precond.jacobi())  # one defect-correction pass
    gt = _np(s.compliance_grad(u0))
    # A is linear in E: A(E +- h) u0 = A u0 +- h B_e u0, with B_e u0 = (A(E + E_e 1_e) - A) u0 / E_e
    # formed once, keeps the rounding of the assembled entries out of the h-sized defect
    r0 = b0 - A0.mv(u0)
    Au0 = A0.mv(u0)
    rng = np.random.default_rng(5)
    for el in rng.choice(T, 4, replace=False):
        h = 1e-6 * E[el]
        EE = torch.tensor(E, device="cuda")
        EE[el] *= 2.0
        Bu = (s.evaluate({"E": EE})[0].mv(u0) - Au0) / E[el]
        ds = []
        for sg in (1, -1):
            EE = torch.tensor(E, device="cuda")
            EE[el] += sg * h
            A2, _ = s.evaluate({"E": EE})
            defect = r0 - (sg * h) * Bu
            d1 = fine(A2, defect, M=precond.jacobi())
            ds.append(d1 + fine(A2, defect - A2.mv(d1), M=precond.jacobi()))
        # b does not depend on E: C+ - C- = b . (d+ - d-)
        fd = float(torch.dot(b0, ds[0] - ds[1])) / (2 * h)
        assert abs(fd - gt[el]) <= 1e-6 * abs(gt[el]), (el, fd, gt[el])


def test_c5_32_elasticity(dev):
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.elasticity import ElasticitySystem

    n = 32
    s = ElasticitySystem(J.kuhn_box(n))
    m = s.mesh
    assert m.n_nodes == 274_625 and m.n_cells == 196_608
    assert s.ndof == 823_875
    assert s.plan.nnz == 69_108_489
    assert int((s.dmask > 0).sum()) == 3 * 65**2
    E = torch.tensor(np.exp(0.3 * np.random.default_rng(0).standard_normal(m.n_cells)), device="cuda")
    A, b = s.evaluate({"E": E})
    assert abs(float(b.sum()) + 1.0) < 1e-12  # unit traction on the unit face
    mixed = solve.cg(tol=1e-10, precision="mixed")
    u = mixed(A, b, M=precond.jacobi())
    f64 = solve.cg(tol=1e-10)
    u64 = f64(A, b, M=precond.jacobi())
    err = _rel(_np(u), _np(u64))
    print(f"\nC5 32^3 P2: mixed {mixed.last_stats}, fp64 {f64.last_stats}, mixed vs fp64 {err:.2e}")
    assert mixed.last_stats["relres"] <= 1e-10  # mixed stops on the true FP64 residual
    assert f64.last_stats["relres_recursive"] <= 1e-10 and f64.last_stats["relres"] <= 2e-10
    assert err <= 1e-8
    g = s.compliance_grad(u)
    assert bool((g <= 0).all())
    # energy identity: C = b.u = u^T A u; sum_e E_e dC/dE_e = -u^T A u on the free dofs
    C = float(torch.dot(b, u))
    assert abs(float(torch.dot(u, A.mv(u))) - C) <= 1e-8 * abs(C)
    assert abs(float(torch.dot(E, g)) + C) <= 1e-8 * abs(C)


# quadratic fields on the face x = 1 and their closed-form integrals over [0,1]^2 in (y, z)
_FACE_FIELDS = (
    (lambda x, y, z: 1.0 + 0 * x, 1.0),
    (lambda x, y, z: z, 0.5),
    (lambda x, y, z: y * y, 1.0 / 3.0),
    (lambda x, y, z: y * z, 0.25),
    (lambda x, y, z: x * z * z - 2.0 * y + 3.0, 1.0 / 3.0 - 1.0 + 3.0),
)


@pytest.mark.parametrize("n", [3, 32])
def test_traction_load_integrates_quadratics_exactly(dev, n):
    # P2 interpolates every quadratic exactly, so the consistent load satisfies
    # sum_i b_i v(x_i) = t int_{x=1} v dS; a vertex/edge swap or a lumped split keeps the total force
    # and fails this (surface term, jno/utils/solver/fem_native.py:1882-1960)
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem

    t = (0.3, -0.7, 1.1)
    s = ElasticitySystem(J.kuhn_box(n), traction=t)
    pts = _np(s.mesh.points)
    b = _np(s.traction_load()).reshape(-1, 3)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    for fn, exact in _FACE_FIELDS:
        assert np.abs(b.T @ fn(x, y, z) - np.asarray(t) * exact).max() <= 1e-13
    if n == 3:  # and node by node against the oracle's load
        p2, c2 = O.promote_p2(*O.kuhn_box(n))
        tris = O.plane_facets_p2(p2, c2, 0, 1.0)
        contrib = O.facet_traction_p2(p2, tris, t)
        bo = np.zeros((p2.shape[0], 3))
        for c in range(3):
            np.add.at(bo[:, c], tris.reshape(-1), contrib[:, :, c].reshape(-1))
        assert np.abs(b - bo).max() <= 1e-15


@pytest.mark.parametrize("n", [2, 3])
def test_vector_traction_recovers_manufactured_displacement(dev, n):
    # tests/test_fem_3d.py:364-382 re-expressed for the P2 elasticity form: nu = 0, E = 1, clamp at
    # x = 0 and t = (1, 0, 0) on x = 1 give u = (x, 0, 0) exactly (sigma = e_x e_x)
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.elasticity import ElasticitySystem

    s = ElasticitySystem(J.kuhn_box(n), nu=0.0, traction=(1.0, 0.0, 0.0))
    A, b = s.evaluate({"E": torch.ones(s.mesh.n_cells, dtype=torch.float64, device="cuda")})
    u = _np(solve.cg(tol=1e-13)(A, b, M=precond.jacobi())).reshape(-1, 3)
    x = _np(s.mesh.points)[:, 0]
    assert np.abs(u - np.stack([x, 0 * x, 0 * x], 1)).max() <= 1e-10
