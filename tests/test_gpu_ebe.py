"""The element-by-element (matrix-free) P2 elasticity operator (the reference's ElementOperator,
jno/utils/solver/element_operator.py:1-178) against the oracle's assembled CSR: SpMV (FP64 to 1e-13,
FP32 to 1e-5), exact diagonal, solves to the oracle's direct solution, geometry classes, and C5 at
full size."""
import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(n, seed, box=None, nu=0.3):
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem

    if box is None:
        mesh, (p, t) = J.kuhn_box(n), O.kuhn_box(n)
    else:
        lo, hi = box
        mesh, (p, t) = J.Shape.box(*lo, *hi).structured(n).domain(), O.kuhn_box(*n, lo, hi)
    p2, c2 = O.promote_p2(p, t)
    E = np.exp(0.3 * np.random.default_rng(seed).standard_normal(c2.shape[0]))
    s = ElasticitySystem(mesh, nu=nu)
    Ao, bo, d = O.elasticity_p2_system(p2, c2, E, nu=nu, load=(0, float(s.load_value)))
    return s, torch.tensor(E, device="cuda"), Ao, bo


@pytest.mark.parametrize("n,box", [(2, None), (3, None), ((4, 2, 2), ((0.0, 0.0, 0.0), (2.0, 1.0, 0.5)))])
def test_ebe_spmv_and_diagonal_match_the_oracle(dev, n, box):
    s, E, Ao, bo = _case(n, 11, box)
    A, b = s.element_operator({"E": E})
    # a dyadic lattice has the 6 Kuhn shapes bitwise; at spacing 1/3 the vertex differences differ in
    # the last bit between cells, and the classes fragment (still exact: one block per bitwise class)
    dyadic = n != 3
    assert (A.geometry.n_classes == 6) == dyadic and A.bcoo is None and A.T is A
    assert np.abs(_np(b) - bo).max() <= 1e-15 * max(1.0, np.abs(bo).max())
    rng = np.random.default_rng(2)
    for _ in range(3):
        x = rng.standard_normal(A.shape[0])
        ref = Ao @ x
        y64 = _np(A.mv(torch.tensor(x, device="cuda")))
        assert _rel(y64, ref) <= 1e-13
        y32 = _np(A.mv(torch.tensor(x, dtype=torch.float32, device="cuda"))).astype(np.float64)
        assert _rel(y32, ref) <= 1e-5
    assert np.abs(_np(A.diag()) - Ao.diagonal()).max() <= 1e-14 * np.abs(Ao.diagonal()).max()
    # assemble() gives the explicit CSR of the same operator
    Ac = A.assemble()
    xt = torch.tensor(rng.standard_normal(A.shape[0]), device="cuda")
    assert _rel(_np(A.mv(xt)), _np(Ac.mv(xt))) <= 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [2, 3])
def test_ebe_solve_matches_the_oracle(dev, precision, n):
    from jno_amd import precond, solve

    s, E, Ao, bo = _case(n, 20 + n)
    A, b = s.element_operator({"E": E})
    sv = solve.cg(tol=1e-10, precision=precision)
    u = sv(A, b, M=precond.jacobi())
    st = sv.last_stats
    assert st["operator"]["format"] == "ebe"
    assert st["relres"] <= (1e-10 if precision == "mixed" else 2e-10), st
    assert _rel(_np(u), O.direct_solve(Ao, bo)) <= 1e-9, st


def test_ebe_every_iteration_mode(dev):
    # the fused modes fall back to the split iteration on EBE operators, as on BSR3
    from jno_amd import precond, solve

    s, E, Ao, bo = _case(3, 31)
    A, b = s.element_operator({"E": E})
    uo = O.direct_solve(Ao, bo)
    base = solve.Tuning.defaults()
    for precision in ("fp64", "mixed"):
        for mode in (0, 1, 3, 4, 5):
            for blk, cap in ((256, 8), (1024, 512)):
                sv = solve.cg(tol=1e-10, precision=precision, tuning=base.with_shape(precision, blk, cap, 0, mode))
                u = sv(A, b, M=precond.jacobi())
                assert _rel(_np(u), uo) <= 1e-9, (precision, mode, blk, cap, sv.last_stats)


def test_ebe_refuses_cells_without_shared_geometry(dev):
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem

    n = 3
    p, t = O.kuhn_box(n)
    pj = p.copy()
    interior = np.ones(p.shape[0], bool)
    interior[O.boundary_nodes(p, t)] = False
    pj[interior] += np.random.default_rng(5).uniform(-0.2, 0.2, (interior.sum(), 3)) / n
    m = J.Mesh(points=torch.tensor(pj, device="cuda"), cells=J.kuhn_box(n).cells, box=((0.0,) * 3, (1.0,) * 3))
    s = ElasticitySystem(m)
    with pytest.raises(ValueError, match="repeated cell geometry"):
        s.element_operator({"E": torch.ones(s.mesh.n_cells, dtype=torch.float64, device="cuda")})


def test_c5_32_ebe_solve(dev):
    # C5 at full size on the EBE operator: the same solution as the assembled BSR3 path
    import time

    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.elasticity import ElasticitySystem

    n = 32
    s = ElasticitySystem(J.kuhn_box(n))
    E = torch.tensor(np.exp(0.3 * np.random.default_rng(0).standard_normal(s.mesh.n_cells)), device="cuda")
    A, b = s.element_operator({"E": E})
    assert A.geometry.n_classes == 6 and A.shape == (823_875, 823_875) and A.structured
    # the half-lattice product against the general element pass + gather at full size
    G, _ = s.element_operator({"E": E}, structured=False)
    xr = torch.randn(A.shape[0], device="cuda", dtype=torch.float32, generator=torch.Generator("cuda").manual_seed(3))
    assert _rel(_np(A.mv(xr)).astype(np.float64), _np(G.mv(xr)).astype(np.float64)) <= 1e-6
    del G
    # the class-table assembly (elimination fused) against the MFMA element matrices + colored scatter
    Ac, _ = s.evaluate({"E": E})
    s.class_assembly = False
    Ae, _ = s.evaluate({"E": E})
    s.class_assembly = True
    assert torch.equal(Ac.col, Ae.col)
    assert float((Ac.vals - Ae.vals).abs().max()) <= 1e-13 * float(Ae.vals.abs().max())
    del Ac, Ae
    mixed = solve.cg(tol=1e-10, precision="mixed")
    u = mixed(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = mixed(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t_ebe = time.perf_counter() - t0
    Ac, bc = s.evaluate({"E": E})
    u64 = solve.cg(tol=1e-10)(Ac, bc, M=precond.jacobi())
    err = _rel(_np(u), _np(u64))
    print(f"\nC5 32^3 EBE mixed: {mixed.last_stats}, {t_ebe * 1e3:.1f} ms; vs assembled FP64 {err:.2e}")
    assert mixed.last_stats["relres"] <= 1e-10 and err <= 1e-8
    # energy identity of the compliance gradient on the EBE solution
    g = s.compliance_grad(u)
    C = float(torch.dot(b, u))
    assert abs(float(torch.dot(E, g)) + C) <= 1e-8 * abs(C)


@pytest.mark.parametrize("n,box", [(1, None), (2, None), (4, None), ((5, 3, 6), ((0.0, 0.0, 0.0), (5.0, 3.0, 6.0))),
                                   ((9, 4, 7), ((0.0, 0.0, 0.0), (4.5, 2.0, 3.5)))])
def test_ebe_structured_box_spmv(dev, n, box):
    # the half-lattice SpMV of a structured Kuhn box (grids that are not multiples of the 4-voxel
    # brick included) against the oracle's CSR and the general element-by-element kernels
    s, E, Ao, bo = _case(n, 7, box)
    A, b = s.element_operator({"E": E})
    G, _ = s.element_operator({"E": E}, structured=False)
    assert A.describe()["structured"] and not G.describe()["structured"]
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.standard_normal(A.shape[0])
        ref = Ao @ x
        xt = torch.tensor(x, dtype=torch.float32, device="cuda")
        y = _np(A.mv(xt)).astype(np.float64)
        yg = _np(G.mv(xt)).astype(np.float64)
        assert _rel(y, ref) <= 1e-5 and _rel(y, yg) <= 1e-5
        # constrained rows return x exactly; repeated products are bitwise equal
        d = np.flatnonzero(_np(s.dmask))
        assert np.array_equal(y[d], x.astype(np.float32)[d].astype(np.float64))
        assert np.array_equal(_np(A.mv(xt)), _np(A.mv(xt)))


def test_ebe_structured_box_solve_matches_general(dev):
    from jno_amd import precond, solve

    s, E, Ao, bo = _case((6, 4, 5), 9, ((0.0, 0.0, 0.0), (1.5, 1.0, 1.25)))
    A, b = s.element_operator({"E": E})
    G, _ = s.element_operator({"E": E}, structured=False)
    assert A.structured
    uo = O.direct_solve(Ao, bo)
    for precision in ("mixed", "fp64"):
        for op in (A, G):
            sv = solve.cg(tol=1e-10, precision=precision)
            u = sv(op, b, M=precond.jacobi())
            assert sv.last_stats["relres"] <= 2e-10 and _rel(_np(u), uo) <= 1e-9, (precision, sv.last_stats)


def test_ebe_box_map_rejects_reordered_cells(dev):
    # two voxels' cells of one Kuhn type swapped: every type keeps one geometry class, but the cells
    # no longer sit where the lattice puts them, so the device check refuses the half-lattice path
    # and the general kernels run the same operator
    import jno_amd as J
    from jno_amd.elasticity import ElasticitySystem, promote_p2

    n = 4
    m1 = J.kuhn_box(n)
    m2 = promote_p2(m1)
    cells = m2.cells.clone()
    t, v1, v2 = 2, 5, 37
    cells[[6 * v1 + t, 6 * v2 + t]] = cells[[6 * v2 + t, 6 * v1 + t]]
    m2s = J.Mesh(points=m2.points, cells=cells.contiguous(), order=2, box=m2.box, grid=m2.grid)
    s = ElasticitySystem(m2s)
    E = torch.tensor(np.exp(0.3 * np.random.default_rng(4).standard_normal(s.mesh.n_cells)), device="cuda")
    A, _ = s.element_operator({"E": E})
    assert A.geometry.n_classes == 6 and not A.structured
    p2, c2 = _np(m2s.points), _np(m2s.cells)
    Ao, _, _ = O.elasticity_p2_system(p2, c2, _np(E), nu=0.3, load=(0, 1.0))
    x = np.random.default_rng(6).standard_normal(A.shape[0])
    assert _rel(_np(A.mv(torch.tensor(x, device="cuda"))), Ao @ x) <= 1e-13
