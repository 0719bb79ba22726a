"""The structured Kuhn-box P1 assembly (jno_assemble_p1_box) against the incidence-gather assembly
(jno_assemble_p1 / jno_assemble_p1_dia) and the oracle.

Bars: bitwise against the incidence-gather kernel on dyadic spacings (every cell of a type has the
same vertex differences there, so both kernels add the same doubles in the same order); 1e-12
relative against the oracle on any spacing."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu

VAL_TOL = 1e-12
SOL_TOL = 1e-9


def _np(t):
    return t.detach().cpu().numpy()


def _bits(t):
    return _np(t).view(np.uint8)


def _box(dims, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    import jno_amd as J

    return J.Shape.box(*lo, *hi).structured(dims).domain()


def _pair(dims, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), **kw):
    """The same problem on the structured-box kernel and on the incidence-gather kernel."""
    import jno_amd as J

    m = _box(dims, lo, hi)
    return m, J.fem_poisson(m, **kw), J.fem_poisson(m, structured=False, **kw)


def test_box_detection(dev):
    import jno_amd as J
    from jno_amd.fem import BoxAssembly

    m = J.kuhn_box(4, 3, 5)
    box = BoxAssembly.of(m)
    assert box is not None and box.grid == (4, 3, 5)
    # Kuhn cubes of a dyadic side: the 7-point couplings only (the other dot products are exact zeros)
    assert BoxAssembly.of(J.kuhn_box(8)).stiffness_offsets == [0, 1, 9, 81]
    # fewer than 2 cells on an axis: the incidence-gather kernel
    assert BoxAssembly.of(J.kuhn_box(1, 4, 4)) is None
    # the same points with the cells' vertices reordered, or another grid claimed: not the box
    cells = m.cells[:, [0, 2, 1, 3]].contiguous()
    assert BoxAssembly.of(J.Mesh(points=m.points, cells=cells, box=m.box, grid=m.grid)) is None
    assert BoxAssembly.of(J.Mesh(points=m.points, cells=m.cells, box=m.box, grid=(4, 5, 3))) is None
    # moving the points in place (torch) checks again and falls back for good
    m2 = J.kuhn_box(4)
    assert BoxAssembly.of(m2) is not None
    m2.points[5, 0] += 1e-3
    assert BoxAssembly.of(m2) is None
    m2.points[5, 0] -= 1e-3
    assert BoxAssembly.of(m2) is None
    m2.invalidate()
    assert BoxAssembly.of(m2) is not None


@pytest.mark.parametrize("dims,hi", [((8, 8, 8), (1.0, 1.0, 1.0)), ((2, 4, 16), (1.0, 2.0, 0.5)),
                                     ((16, 2, 2), (4.0, 0.25, 0.125))])
def test_box_bitwise_on_dyadic_spacings(dev, dims, hi):
    """Dyadic spacings: every value of A, b and the Jacobi inverse equals the incidence-gather
    kernel's bitwise, with a nonzero Dirichlet lift, a nodal load and a constant load."""
    m, sb, sg = _pair(dims, hi=hi)
    p = _np(m.points)
    k = torch.tensor(O.k_lognormal(m.n_cells, 17), device="cuda")
    f = torch.tensor(1.0 + p[:, 0] * p[:, 1] - p[:, 2], device="cuda")
    g = torch.tensor(np.sin(p[:, 0]) + p[:, 2], device="cuda")
    import jno_amd as J

    for kw in ({}, {"f": f, "dirichlet": "left", "g": g}, {"f": -2.5, "dirichlet": "boundary", "g": g},
               {"dirichlet": None}):
        sb = J.fem_poisson(m, **kw)
        sg = J.fem_poisson(m, structured=False, **kw)
        Ab, bb = sb.evaluate({"k": k})
        Ag, bg = sg.evaluate({"k": k})
        assert np.array_equal(_bits(Ab.vals), _bits(Ag.vals)), kw
        assert np.array_equal(_bits(bb), _bits(bg)), kw
        assert np.array_equal(_bits(Ab.dinv), _bits(Ag.dinv)), kw


@pytest.mark.parametrize("dims,lo,hi", [((3, 5, 7), (-1.0, 0.5, 2.0), (3.0, 1.7, 2.3)),
                                        ((2, 2, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
                                        ((7, 3, 2), (0.1, 0.2, 0.3), (0.8, 1.9, 0.7))])
def test_box_vs_oracle(dev, dims, lo, hi):
    """Any spacing (voxel 0's geometry stands for its type): A and b within 1e-12 of the oracle's
    assembly, nonzero Dirichlet data on a named face and a nodal load."""
    m, sb, sg = _pair(dims, lo, hi)
    p, t = O.kuhn_box(*dims, lo, hi)
    assert np.array_equal(_np(m.points), p)
    k = O.k_lognormal(t.shape[0], 5)
    f = 1.0 + p[:, 0] * p[:, 1]
    g = np.cos(p[:, 1]) - p[:, 0]
    import jno_amd as J

    s = J.fem_poisson(m, f=torch.tensor(f, device="cuda"), dirichlet="bottom", g=torch.tensor(g, device="cuda"))
    A, b = s.evaluate({"k": torch.tensor(k, device="cuda")})
    N = p.shape[0]
    rp, col, vals = O.assemble(O.p1_stiffness(p, t, k), t, 1, N)
    dofs = O.boundary_nodes(p, t, "bottom", lo, hi)
    Ao, bo = O.dirichlet_eliminate(O.csr_matrix(rp, col, vals, N), O.p1_load(p, t, f, N), dofs, g[dofs])
    Ao.sort_indices()
    assert np.array_equal(_np(A.col), Ao.indices)
    assert np.abs(_np(A.vals) - Ao.data).max() <= VAL_TOL * np.abs(Ao.data).max()
    assert np.abs(_np(b) - bo).max() <= VAL_TOL * np.abs(bo).max()
    d = Ao.diagonal()
    assert np.abs(_np(A.dinv) - 1.0 / d).max() <= VAL_TOL * np.abs(1.0 / d).max()


@pytest.mark.parametrize("dims,hi", [((8, 4, 8), (1.0, 1.0, 1.0)), ((3, 5, 4), (0.7, 1.3, 2.9))])
def test_box_mass(dev, dims, hi):
    m, sb, sg = _pair(dims, hi=hi)
    Mb, Mg = sb.mass(), sg.mass()
    p, t = O.kuhn_box(*dims, (0.0, 0.0, 0.0), hi)
    Mo = O.mass_matrix_p1(p, t)
    Mo.sort_indices()
    assert np.abs(_np(Mb.vals) - Mo.data).max() <= VAL_TOL * Mo.data.max()
    if hi == (1.0, 1.0, 1.0):
        assert np.array_equal(_bits(Mb.vals), _bits(Mg.vals))


@pytest.mark.parametrize("dims,hi", [((16, 8, 8), (1.0, 1.0, 1.0)), ((10, 10, 10), (1.0, 1.0, 1.0)),
                                     ((5, 7, 6), (1.1, 0.9, 1.7))])
def test_box_writes_the_dia_layout(dev, dims, hi):
    """Once a solve has picked DIA for the plan, the box kernel writes the layout in the same pass
    with no flag read back: FP64 and FP32 diagonals bitwise equal to the incidence-gather kernel's
    checked fill on dyadic spacings (and to a fill from the box kernel's own CSR on any), and the
    solve matches the oracle's direct solution."""
    from jno_amd import precond, solve
    from jno_amd.operator import SolverOperator

    m, sb, sg = _pair(dims, hi=hi)
    k = torch.tensor(O.k_lognormal(m.n_cells, 8), device="cuda")
    sv = solve.cg(tol=1e-10, precision="mixed")
    A0, b0 = sb.evaluate({"k": k})
    sv(A0, b0, M=precond.jacobi())
    Ab, bb = sb.evaluate({"k": 2.0 * k})
    Ag, bg = sg.evaluate({"k": 2.0 * k})
    dyadic = hi == (1.0, 1.0, 1.0) and all(d & (d - 1) == 0 for d in dims)
    fb = Ab._cache.get(("solver_op", "auto", True))
    fg = Ag._cache.get(("solver_op", "auto", True))
    assert fb is not None and fb.format == "dia"
    ref = SolverOperator(Ab, fmt="dia")
    assert fb.offsets == ref.offsets
    n = Ab.shape[0]
    ld = (n + 63) // 64 * 64
    for i in range(2):
        vb = fb._keep[i].reshape(fb.ndiag, ld)[:, :n].contiguous()
        vr = ref._keep[i].reshape(ref.ndiag, ld)[:, :n].contiguous()
        assert np.array_equal(_bits(vb), _bits(vr))
        if dyadic:
            assert fg is not None and fg.offsets == fb.offsets
            vg = fg._keep[i].reshape(fg.ndiag, ld)[:, :n].contiguous()
            assert np.array_equal(_bits(vb), _bits(vg))
    u = sv(Ab, bb, M=precond.jacobi())
    p, t = O.kuhn_box(*dims, (0.0, 0.0, 0.0), hi)
    Ao, bo, _ = O.poisson_p1_system(p, t, 2.0 * _np(k))
    ue = O.direct_solve(Ao, bo)
    assert float(np.linalg.norm(_np(u) - ue) / np.linalg.norm(ue)) <= SOL_TOL


def test_box_entry_rejects_bad_arguments(dev):
    from jno_amd import _lib

    m = _box((4, 4, 4))
    plan = m.plan(1)
    geo = (C.c_double * (6 * 17))()
    vals = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    offs = (C.c_int32 * _lib.JNO_DIA_MAX)(0)
    lib = _lib.load()
    st = _lib.stream_ptr()
    args = (_lib.ptr(plan.rowptr), None, 0.0, None, None, None, _lib.ptr(vals), None, None, 0, offs, None, None, st)
    assert lib.jno_assemble_p1_box(0, 4, 4, 4, geo, *args) == _lib.JNO_EINVAL  # stiffness without k
    assert lib.jno_assemble_p1_box(2, 4, 4, 4, geo, *args) == _lib.JNO_EINVAL
    assert lib.jno_assemble_p1_box(1, 1, 4, 4, geo, *args) == _lib.JNO_EINVAL
    ws = torch.empty(64, dtype=torch.uint8, device="cuda")
    assert lib.jno_box_geometry(4, 4, 4, _lib.ptr(m.points), geo, _lib.ptr(ws), 64, st) == _lib.JNO_EWORKSPACE


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 4, 5), (7, 2, 3), (16, 16, 16), (33, 17, 9)])
def test_box_plan_and_mask_equal_the_sorted_ones(dev, dims):
    """The closed-form plan of a Kuhn box (rowptr, col, the incidence and its offsets) and its
    surface mask equal, element for element, what the sort-based builder and the boundary-facet
    mask produce for the same mesh; the pattern also equals the oracle's."""
    import jno_amd as J
    from jno_amd.fem import Plan, plan_for

    m = J.kuhn_box(*dims)
    plan = plan_for(m, 1)
    ref = Plan.build(m, 1)
    assert (plan.nnz, plan.max_row, plan.ndof) == (ref.nnz, ref.max_row, ref.ndof)
    for name in ("rowptr", "col", "inc_ptr", "inc"):
        assert torch.equal(getattr(plan, name), getattr(ref, name)), name
    if np.prod(dims) <= 300:
        p, t = O.kuhn_box(*dims)
        rp, col, _, nse = O.pattern_struct(t, 1)
        assert plan.nnz == nse and np.array_equal(_np(plan.rowptr), rp) and np.array_equal(_np(plan.col), col)
    mask = m.boundary_mask()
    m2 = J.Mesh(points=m.points.clone(), cells=m.cells.clone())  # no grid: the facet-based mask
    assert torch.equal(mask, m2.boundary_mask())
    for face in ("left", "top", "back"):
        assert torch.equal(m.boundary_mask(face), J.Mesh(points=m.points, cells=m.cells, box=m.box).boundary_mask(face))
