"""Parity of the HIP P1 path (mesh, pattern, assembly, Dirichlet, PCG, adjoint) with the oracle.

Bars: bit-exact for meshes, masks and the CSR index structure; 1e-12 relative for assembled values;
<= 1e-9 relative L2 for solutions at rtol 1e-10 (the FP64 oracle is scipy's direct solve)."""
import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu

SOL_TOL = 1e-9  # relative L2 vs the FP64 oracle at rtol 1e-10
VAL_TOL = 1e-12  # relative, assembled values


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (5, 5, 5), (16, 16, 16)])
def test_mesh_bit_exact(dev, dims):
    import jno_amd as J

    m = J.kuhn_box(*dims)
    p, t = O.kuhn_box(*dims)
    assert np.array_equal(_np(m.points), p)
    assert np.array_equal(_np(m.cells), t)


def test_mesh_non_unit_box(dev):
    import jno_amd as J

    lo, hi = (-1.0, 0.5, 2.0), (3.0, 1.7, 2.3)
    m = J.Shape.box(*lo, *hi).structured((3, 5, 7)).domain()
    p, t = O.kuhn_box(3, 5, 7, lo, hi)
    assert np.array_equal(_np(m.points), p) and np.array_equal(_np(m.cells), t)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16])
def test_pattern_bit_exact(dev, n):
    import jno_amd as J

    m = J.kuhn_box(n)
    plan = m.plan(1)
    p, t = O.kuhn_box(n)
    rp, col, _, nse = O.pattern_struct(t, 1)
    assert plan.nnz == nse
    assert np.array_equal(_np(plan.rowptr), rp)
    assert np.array_equal(_np(plan.col), col)
    # incidence: ascending (element*4 + local) per node, covering every cell slot once
    ip, inc = _np(plan.inc_ptr), _np(plan.inc)
    assert ip[-1] == 4 * t.shape[0]
    for r in range(0, p.shape[0], max(1, p.shape[0] // 50)):
        seg = inc[ip[r]: ip[r + 1]]
        assert np.all(np.diff(seg) > 0)
        assert np.all(t.reshape(-1)[seg] == r)


def test_pattern_vec3_bit_exact(dev):
    import jno_amd as J

    m = J.kuhn_box(2)
    plan = m.plan(3)
    p, t = O.kuhn_box(2)
    rp, col, _, _ = O.pattern_struct(t, 3)
    assert np.array_equal(_np(plan.rowptr), rp) and np.array_equal(_np(plan.col), col)


def test_pattern_matches_golden(dev):
    import jno_amd as J

    g = np.load("tests/golden/p1_patterns.npz")
    for n in (2, 4, 16):
        plan = J.kuhn_box(n).plan(1)
        assert np.array_equal(_np(plan.rowptr), g[f"rowptr_{n}"])
        assert np.array_equal(_np(plan.col), g[f"col_{n}"])


@pytest.mark.parametrize("n", [2, 5, 16])
def test_boundary_masks(dev, n):
    import jno_amd as J

    m = J.kuhn_box(n)
    p, t = O.kuhn_box(n)
    assert np.array_equal(np.flatnonzero(_np(m.boundary_mask())), O.boundary_nodes(p, t))
    for face in ("left", "right", "bottom", "top", "front", "back"):
        assert np.array_equal(np.flatnonzero(_np(m.boundary_mask(face))), O.boundary_nodes(p, t, face))


def _system(n, k, **kw):
    import jno_amd as J

    m = J.kuhn_box(n)
    s = J.fem_poisson(m, **kw)
    A, b = s.evaluate({"k": torch.as_tensor(k, device="cuda")})
    return m, s, A, b


def test_assembly_c1_values(dev):
    n = 16
    p, t = O.kuhn_box(n)
    k = O.k_sin(p, t)
    m, s, A, b = _system(n, k)
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    Ao.sort_indices()
    assert np.array_equal(_np(A.col), Ao.indices) and np.array_equal(_np(A.rowptr), Ao.indptr)
    v = _np(A.vals)
    assert np.abs(v - Ao.data).max() <= VAL_TOL * np.abs(Ao.data).max()
    assert np.abs(_np(b) - bo).max() <= VAL_TOL * np.abs(bo).max()
    assert np.abs(_np(A.dinv) - O.jacobi_inv(Ao)).max() <= VAL_TOL * np.abs(O.jacobi_inv(Ao)).max()
    # exact-zero couplings on the dyadic Kuhn mesh: the zero-dropped operator is the oracle's P_eager
    Ae = A.drop_zeros()
    assert Ae.nnz == O.compress_eager(Ao).nnz == 23_813
    # symmetric, bitwise
    D = _np(A.dense())
    assert np.array_equal(D, D.T)


def test_assembly_nonzero_dirichlet_and_nodal_load(dev):
    import jno_amd as J

    n = 6
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 11)
    f = 1.0 + p[:, 0] * p[:, 1]
    g = np.sin(p[:, 0]) + p[:, 2]
    m = J.kuhn_box(n)
    s = J.fem_poisson(m, f=torch.tensor(f, device="cuda"), dirichlet="left", g=torch.tensor(g, device="cuda"))
    A, b = s.evaluate({"k": torch.tensor(k, device="cuda")})
    N = p.shape[0]
    rp, col, vals = O.assemble(O.p1_stiffness(p, t, k), t, 1, N)
    dofs = O.boundary_nodes(p, t, "left")
    Ao, bo = O.dirichlet_eliminate(O.csr_matrix(rp, col, vals, N), O.p1_load(p, t, f, N), dofs, g[dofs])
    assert np.abs(_np(A.vals) - Ao.data).max() <= VAL_TOL * np.abs(Ao.data).max()
    assert np.abs(_np(b) - bo).max() <= VAL_TOL * np.abs(bo).max()


def test_assembly_fan_mesh(dev):
    """Assembly on a mesh off the lattice: a fan whose centre has 40 incident tets of both
    orientations (ring of 20 points in z = 0, apexes above and below), against the oracle."""
    import jno_amd as J

    ring = [(np.cos(2 * np.pi * i / 20), np.sin(2 * np.pi * i / 20), 0.0) for i in range(20)]
    pts = np.array([(0.0, 0.0, 0.0)] + ring + [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0)])
    cells = []
    for i in range(20):
        a, b = 1 + i, 1 + (i + 1) % 20
        cells += [(0, a, b, 21), (0, b, a, 22)]
    cells = np.array(cells, dtype=np.int32)
    fan = J.Mesh(points=torch.tensor(pts, device="cuda"), cells=torch.tensor(cells, device="cuda"))
    k = O.k_lognormal(cells.shape[0], 22)
    A, b = J.fem_poisson(fan, dirichlet=None).evaluate({"k": torch.tensor(k, device="cuda")})
    N = pts.shape[0]
    Ao = O.csr_matrix(*O.assemble(O.p1_stiffness(pts, cells, k), cells, 1, N), N)
    Ao.sort_indices()
    assert np.array_equal(_np(A.col), Ao.indices) and np.array_equal(_np(A.rowptr), Ao.indptr)
    assert np.abs(_np(A.vals) - Ao.data).max() <= VAL_TOL * np.abs(Ao.data).max()
    assert np.abs(_np(b) - O.p1_load(pts, cells, 1.0, N)).max() <= VAL_TOL * np.abs(_np(b)).max()


def test_mass_matrix(dev):
    n = 5
    p, t = O.kuhn_box(n)
    m, s, A, b = _system(n, np.ones(6 * n**3))
    M = s.mass()
    Mo = O.mass_matrix_p1(p, t)
    Mo.sort_indices()
    assert np.abs(_np(M.vals) - Mo.data).max() <= VAL_TOL * Mo.data.max()


def test_spmv_and_generic_dirichlet(dev):
    import jno_amd as J
    from jno_amd import _lib

    n = 7
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 5)
    m = J.kuhn_box(n)
    s = J.fem_poisson(m, dirichlet=None)
    A, b = s.evaluate({"k": torch.tensor(k, device="cuda")})
    Ao = O.csr_matrix(*O.assemble(O.p1_stiffness(p, t, k), t, 1, p.shape[0]), p.shape[0])
    x = np.random.default_rng(0).standard_normal(p.shape[0])
    assert _rel(_np(A.mv(torch.tensor(x, device="cuda"))), Ao @ x) < 1e-13
    # jno_dirichlet_eliminate on the assembled CSR == fused elimination == oracle
    mask = m.boundary_mask()
    g = torch.tensor(np.cos(p[:, 1]), device="cuda")
    _lib.call("jno_dirichlet_eliminate", A.shape[0], _lib.ptr(A.rowptr), _lib.ptr(A.col), _lib.ptr(A.vals), _lib.ptr(b),
              _lib.ptr(mask), _lib.ptr(g), _lib.stream_ptr())
    dofs = O.boundary_nodes(p, t)
    A2, b2 = O.dirichlet_eliminate(Ao, O.p1_load(p, t, 1.0, p.shape[0]), dofs, np.cos(p[dofs, 1]))
    assert np.abs(_np(A.vals) - A2.data).max() <= VAL_TOL * np.abs(A2.data).max()
    assert np.abs(_np(b) - b2).max() <= VAL_TOL * np.abs(b2).max()


@pytest.mark.parametrize("fmt", ["csr", "dia", "auto"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_c1(dev, fmt, precision):
    from jno_amd import precond, solve

    n = 16
    p, t = O.kuhn_box(n)
    k = O.k_sin(p, t)
    m, s, A, b = _system(n, k)
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    xd = O.direct_solve(Ao, bo)
    solver = solve.cg(tol=1e-10, precision=precision, format=fmt)
    u = solver(A, b, M=precond.jacobi())
    st = solver.last_stats
    assert _rel(_np(u), xd) <= SOL_TOL, st
    assert st["relres"] <= 1e-10, st
    if precision == "fp64":
        # jax cg semantics: the oracle's restated iteration stops after 64 iterations
        assert abs(st["iterations"] - 64) <= 1, st
    gold = np.load("tests/golden/c1_solution.npz")
    assert _rel(_np(u), gold["u"]) <= SOL_TOL


def test_solve_semantics(dev):
    from jno_amd import precond, solve

    n = 8
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 9)
    m, s, A, b = _system(n, k)
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    # b = 0: zero iterations, x = 0
    sv = solve.cg(tol=1e-10)
    x = sv(A, torch.zeros_like(b), M=precond.jacobi())
    assert sv.last_stats["iterations"] == 0 and not _np(x).any()
    # the iteration count follows the oracle's restated jax cg
    xo, ito = O.pcg(Ao, bo, O.jacobi_inv(Ao), tol=1e-10)
    sv(A, b, M=precond.jacobi())
    assert abs(sv.last_stats["iterations"] - ito) <= 1
    # x0 = solution: converged at once
    x = sv(A, b, M=precond.jacobi(), x0=torch.tensor(O.direct_solve(Ao, bo), device="cuda"))
    assert sv.last_stats["iterations"] <= 1
    # maxiter too small: the 1e-4 residual guard raises like _maybe_residual_check
    with pytest.raises(RuntimeError, match="did not solve the system"):
        solve.cg(tol=1e-10, maxiter=2)(A, b, M=precond.jacobi())
    # M = None (identity) still converges
    x = solve.cg(tol=1e-10)(A, b)
    assert _rel(_np(x), O.direct_solve(Ao, bo)) <= SOL_TOL
    # a non-diagonal applier is refused with a remedy
    with pytest.raises(TypeError, match="diagonal"):
        solve.cg()(A, b, M=lambda v: v)


def test_solve_nonfinite_coefficient_is_detected(dev):
    from jno_amd import precond, solve

    n = 4
    k = np.ones(6 * n**3)
    k[7] = np.nan
    m, s, A, b = _system(n, k)
    with pytest.raises(RuntimeError, match="did not solve"):
        solve.cg(tol=1e-10)(A, b, M=precond.jacobi())


def test_operator_layouts_agree(dev):
    from jno_amd.operator import SolverOperator

    # h = 1/10 is not dyadic: the structural zeros still come out exactly 0.0 (contraction-free
    # geometry), so the nonzero pattern is the 7-point stencil and DIA keeps 4 diagonals
    n = 10
    p, t = O.kuhn_box(n)
    m, s, A, b = _system(n, O.k_lognormal(t.shape[0], 2))
    x = torch.randn(A.shape[0], dtype=torch.float64, device="cuda")
    ref = _np(A.mv(x))
    dia = SolverOperator(A, fmt="dia")
    assert dia.ndiag == 4 and dia.offsets == [0, 1, n + 1, (n + 1) ** 2]
    D = _np(A.dense())
    assert np.array_equal(D, D.T)
    assert _rel(_np(dia.mv(x)), ref) < 1e-14
    for L in (1, 2, 4, 8, 16, 32, 64):
        op = SolverOperator(A, fmt="csr", lanes=L)
        assert _rel(_np(op.mv(x)), ref) < 1e-14
        assert _rel(_np(op.mv(x.float())).astype(np.float64), ref) < 1e-6


def test_dia_offsets_reused_with_checked_fill(dev, monkeypatch):
    """A re-assembly on the same plan reuses the DIA offsets without a new analysis; values that the
    cached layout cannot reproduce (an asymmetric entry, an entry off the stored diagonals) fail the
    device check and the full analysis runs."""
    from jno_amd.fem import CSR
    from jno_amd.operator import SolverOperator

    n = 9
    m, s, A, b = _system(n, O.k_lognormal(O.kuhn_box(n)[1].shape[0], 3))
    op1 = SolverOperator(A, fmt="dia")
    A2, _ = s.evaluate({"k": torch.tensor(O.k_lognormal(m.n_cells, 4), device="cuda")})
    assert A2.rowptr.data_ptr() == A.rowptr.data_ptr() and A2.col.data_ptr() == A.col.data_ptr()
    calls = []
    analyze = SolverOperator.analyze
    monkeypatch.setattr(SolverOperator, "analyze", staticmethod(lambda M: calls.append(1) or analyze(M)))
    op2 = SolverOperator(A2, fmt="dia")
    assert not calls and op2.format == "dia" and op2.offsets == op1.offsets
    x = torch.randn(A.shape[0], dtype=torch.float64, device="cuda")
    assert _rel(_np(op2.mv(x)), _np(A2.mv(x))) < 1e-14
    assert _rel(_np(op2.mv(x.float())).astype(np.float64), _np(A2.mv(x))) < 1e-6

    rp, col, v = _np(A2.rowptr), _np(A2.col), _np(A2.vals).copy()
    rows = np.repeat(np.arange(A2.shape[0]), np.diff(rp))

    def mirror(j):
        r, c = rows[j], col[j]
        return rp[c] + int(np.nonzero(col[rp[c]:rp[c + 1]] == r)[0][0])

    # an asymmetric lower entry: no symmetric layout, the format choice falls back to CSR
    j = int(np.nonzero((col < rows) & (v != 0.0))[0][0])
    va = v.copy()
    va[j] *= 1.5
    A3 = CSR(A2.rowptr, A2.col, torch.tensor(va, device="cuda"), A2.shape)
    op3 = SolverOperator(A3, fmt="auto")
    assert calls and op3.format == "csr"
    assert _rel(_np(op3.mv(x)), _np(A3.mv(x))) < 1e-14
    # a symmetric pair on a structural-zero diagonal: DIA applies again, with more diagonals
    calls.clear()
    off_set = (col > rows) & (v == 0.0) & ~np.isin(col - rows, op1.offsets)
    j = int(np.nonzero(off_set)[0][0])
    vz = v.copy()
    vz[j] = vz[mirror(j)] = -0.125
    A4 = CSR(A2.rowptr, A2.col, torch.tensor(vz, device="cuda"), A2.shape)
    op4 = SolverOperator(A4, fmt="dia")
    assert calls and op4.ndiag == 5
    assert _rel(_np(op4.mv(x)), _np(A4.mv(x))) < 1e-14


def test_assembly_writes_dia_layout(dev):
    """Once a solve has picked DIA for a plan, re-assembly writes the layout in the same pass: the
    diagonals equal bitwise those filled from the CSR.  Moving the mesh points off the lattice
    (same plan) puts nonzeros off the stored diagonals: the fused check refuses the layout and the
    solve still matches the oracle."""
    import jno_amd as J
    from jno_amd import precond, solve
    from jno_amd.operator import SolverOperator

    n = 10
    p, t = O.kuhn_box(n)
    m = J.kuhn_box(n)
    s = J.fem_poisson(m)
    k = torch.tensor(O.k_lognormal(t.shape[0], 31), device="cuda")
    A0, b0 = s.evaluate({"k": k})
    sv = solve.cg(tol=1e-10, precision="mixed")
    sv(A0, b0, M=precond.jacobi())  # picks (and remembers) the DIA offsets of this plan
    A, b = s.evaluate({"k": k * 1.5})
    fused = A._cache.get(("solver_op", "auto", True))
    assert fused is not None and fused.format == "dia"
    ref = SolverOperator(A, fmt="dia")
    ld = (A.shape[0] + 63) // 64 * 64
    assert fused.offsets == ref.offsets
    for f_buf, r_buf in ((fused._keep[0], ref._keep[0]), (fused._keep[1], ref._keep[1])):
        fb = f_buf.reshape(fused.ndiag, ld)[:, : A.shape[0]].contiguous()
        rb = r_buf.reshape(ref.ndiag, ld)[:, : A.shape[0]].contiguous()
        assert np.array_equal(_np(fb).view(np.uint8), _np(rb).view(np.uint8))
    u = sv(A, b, M=precond.jacobi())
    Ao, bo, _ = O.poisson_p1_system(p, t, 1.5 * _np(k))
    assert _rel(_np(u), O.direct_solve(Ao, bo)) <= SOL_TOL
    # off the lattice: interior points moved by up to 0.2 h, same cells and plan
    rng = np.random.default_rng(3)
    interior = np.ones(p.shape[0], bool)
    interior[O.boundary_nodes(p, t)] = False
    pj = p.copy()
    pj[interior] += rng.uniform(-0.2, 0.2, (interior.sum(), 3)) / n
    m.points.copy_(torch.tensor(pj, device="cuda"))
    A2, b2 = s.evaluate({"k": k})
    assert ("solver_op", "auto", True) not in A2._cache
    u2 = sv(A2, b2, M=precond.jacobi())
    Ao2, bo2, _ = O.poisson_p1_system(pj, t, _np(k))
    assert _rel(_np(u2), O.direct_solve(Ao2, bo2)) <= SOL_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_iteration_modes_and_launch_shapes(dev, precision):
    """Every launch shape of the PCG iteration kernels solves to the oracle's direct solution:
    iteration mode K1 K2 / K0 KS K2 / KA KB (single reduction) / K0 KS K2, K1 K2 and one kernel on
    the Jacobi-scaled FP32 operator, SpMV variant (2 rows per thread, 16-byte tiles, 4 rows per thread,
    p staged in LDS), several workgroup sizes and grid caps, DIA and CSR."""
    from jno_amd import precond, solve

    n = 11  # non-dyadic, ragged row ranges per workgroup
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 3)
    m, s, A, b = _system(n, k)
    Ao, bo, _ = O.poisson_p1_system(p, t, k)
    uo = O.direct_solve(Ao, bo)
    base = solve.Tuning.defaults()
    for fmt in ("dia", "csr"):
        for mode in (0, 1, 2, 3, 4, 5):
            for vec in (0, 1, 2, 3):
                for blk, cap in ((256, 8), (1024, 64), (512, 4096)):
                    tun = base.with_shape(precision, blk, cap, vec, mode)
                    sv = solve.cg(tol=1e-10, precision=precision, format=fmt, tuning=tun)
                    u = sv(A, b, M=precond.jacobi())
                    st = sv.last_stats
                    # jax semantics stop FP64 on the recursive residual; mixed stops on the true one
                    assert st["relres"] <= (1e-10 if precision == "mixed" else 2e-10), (fmt, mode, vec, blk, cap, st)
                    assert _rel(_np(u), uo) <= SOL_TOL, (fmt, mode, vec, blk, cap, st)


@pytest.mark.parametrize("dims,hi,ndiag", [((5, 9, 40), (5 / 8, 9 / 8, 5.0), 4),
                                           ((30, 7, 3), (30 / 16, 7 / 16, 3 / 16), 4),
                                           ((9, 20, 300), (9 / 32, 20 / 32, 300 / 32), 4),
                                           ((5, 9, 40), (1.0, 1.0, 1.0), 4)])
def test_one_kernel_non_cubic_boxes(dev, dims, hi, ndiag):
    """The one-kernel FP32 iterations on non-cubic Kuhn boxes: the operator has 4 DIA offsets (1,
    nz+1, (ny+1)(nz+1)), not those of a cube. The cases:
    * 2 near diagonals and a far one (5 x 9 x 40: offsets 1, 41, 410);
    * all 3 near (30 x 7 x 3: 1, 4, 32);
    * the first far offset past the LDS halo (9 x 20 x 300: 1, 301, 6321). The register-resident
      kernel does not take it, so its launch shape falls back to the LDS-staged kernel;
    * unequal spacings (5 x 9 x 40 on the unit cube: 1, 41, 410). The Kuhn subdivision of a box
      gives the 7-point stencil for any spacings, and the contraction-free assembly keeps the
      other couplings exactly 0 (the oracle's differ from 0 by round-off, |a| < 2e-16 against
      couplings of 0.03 to 2.6).
    Each variant (mode 5 with k1_vector 0, 1, 2) and the default shape solve to the oracle's direct
    solution."""
    import jno_amd as J
    from jno_amd import precond, solve

    nx, ny, nz = dims
    p, t = O.kuhn_box(nx, ny, nz, hi=hi)
    k = O.k_lognormal(t.shape[0], 11)
    m = J.kuhn_box(nx, ny, nz, hi=hi)
    assert np.array_equal(_np(m.cells), t)
    assert np.array_equal(_np(m.points), p)
    A, b = J.fem_poisson(m).evaluate({"k": torch.as_tensor(k, device="cuda")})
    Ao, bo, _ = O.poisson_p1_system(p, t, k)
    uo = O.direct_solve(Ao, bo)
    base = solve.Tuning.defaults()
    for shape in (None, (512, 512, 0, 5), (256, 768, 1, 5), (256, 768, 2, 5), (256, 768, 3, 5), (512, 64, 2, 5), (512, 64, 3, 5)):
        tun = base.with_shape("mixed", *shape) if shape is not None else None
        sv = solve.cg(tol=1e-10, precision="mixed", replace_factor=0.01, tuning=tun)
        u = sv(A, b, M=precond.jacobi())
        st = sv.last_stats
        assert st["operator"]["format"] == "dia" and st["operator"]["ndiag"] == ndiag, st
        assert st["relres"] <= 1e-10, (dims, shape, st)
        assert _rel(_np(u), uo) <= SOL_TOL, (dims, shape, st)


@pytest.mark.parametrize("n", [17, 40])
def test_lds_staged_k1(dev, n):
    """The LDS-staged kernels on the Jacobi-scaled operator (K1 of mode 4 with k1_vector 3, and the
    one-kernel iteration, mode 5): offsets beyond the LDS halo (256 rows) recomputed from the
    iteration vectors (n = 17: offset 324), several LDS tiles per workgroup (n = 40 with 8
    workgroups: 8616 rows each), residual replacements (replace_factor 0.1 and 0.01).  Reference: an
    FP64 solve at rtol 1e-13."""
    from jno_amd import precond, solve

    k = O.k_lognormal(6 * n**3, 5)
    m, s, A, b = _system(n, k)
    ref = _np(solve.cg(tol=1e-13, precision="fp64")(A, b, M=precond.jacobi()))
    base = solve.Tuning.defaults()
    its = {}
    # (3, 4): K1 with p staged in LDS; (0, 5): one kernel per iteration, LDS-staged; (1, 5) and
    # (2, 5): one kernel per iteration with register-resident 4-row quads, far neighbours loaded
    # before or after the head (256 and 512 threads; 1024 falls back to the LDS-staged kernel);
    # (3, 5): the same with pipelined tiles (several tiles per workgroup at the small grid caps)
    for vec, mode in ((3, 4), (0, 5), (1, 5), (2, 5), (3, 5)):
        for blk, cap in ((1024, 8), (256, 64), (1024, 512), (512, 8), (256, 4096)):
            for rf in (0.1, 0.01):
                sv = solve.cg(tol=1e-10, precision="mixed", replace_factor=rf,
                              tuning=base.with_shape("mixed", blk, cap, vec, mode))
                u = sv(A, b, M=precond.jacobi())
                st = sv.last_stats
                assert st["relres"] <= 1e-10, (vec, mode, blk, cap, rf, st)
                assert _rel(_np(u), ref) <= SOL_TOL, (vec, mode, blk, cap, rf, st)
                assert st["refinements"] >= 2, st
                its.setdefault((vec, mode, rf), []).append(st["iterations"])
    # the one-kernel iteration keeps the search direction across replacements: about the iteration
    # count of the two-kernel one (plus one rerun launch per replacement)
    for rf in (0.1, 0.01):
        for vec in (0, 1, 2):
            assert max(its[(vec, 5, rf)]) <= 1.1 * min(its[(3, 4, rf)]) + 12, its


@pytest.mark.parametrize("n,cap", [(24, 768), (40, 64), (13, 8)])
def test_x_every_other_launch_is_bitwise_eager(dev, n, cap):
    """The pipelined one-kernel iteration (k1_vector 3) updates x every other launch: even launches
    defer their term, odd ones add both, and a stopping odd launch adds the deferred one.  Its
    products and sums are those of the eager update (k1_vector 2, x updated by every launch) in the
    same order, so the solutions are bitwise equal, with the same launch and replacement counts.
    Each solve stops for several replacements, at launches of either parity; the small grid caps give
    several tiles per workgroup (the stop's flush pass walks them all)."""
    from jno_amd import precond, solve

    k = O.k_lognormal(6 * n**3, 21)
    m, s, A, b = _system(n, k)
    base = solve.Tuning.defaults()
    for rf in (0.3, 0.1, 0.01):
        out = {}
        for vec in (2, 3):
            sv = solve.cg(tol=1e-10, precision="mixed", replace_factor=rf,
                          tuning=base.with_shape("mixed", 256, cap, vec, 5))
            u = sv(A, b, M=precond.jacobi())
            out[vec] = (u, dict(sv.last_stats))
        (u2, s2), (u3, s3) = out[2], out[3]
        assert s3["relres"] <= 1e-10, s3
        assert (s2["iterations"], s2["refinements"]) == (s3["iterations"], s3["refinements"]), (s2, s3)
        assert s3["refinements"] >= 2, s3
        assert torch.equal(u2, u3), (n, cap, rf, float((u2 - u3).abs().max()))


def test_adjoint_gradient_vs_oracle_and_fd(dev):
    import jno_amd as J
    from jno_amd import adjoint, precond, solve

    n = 8
    p, t = O.kuhn_box(n)
    T = t.shape[0]
    k = O.k_lognormal(T, 0)
    kstar = O.k_lognormal(T, 1)
    m = J.kuhn_box(n)
    s = J.fem_poisson(m)
    M = s.mass()
    sv = solve.cg(tol=1e-12)
    Astar, bstar = s.evaluate({"k": torch.tensor(kstar, device="cuda")})
    ustar = sv(Astar, bstar, M=precond.jacobi())
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    u = adjoint.solve(s, kt, sv)
    e = u - ustar
    L = 0.5 * torch.dot(e, M.mv(e))
    L.backward()
    grad = _np(kt.grad)
    # oracle: direct solves + closed-form contraction
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    Mo = O.mass_matrix_p1(p, t)
    uo = O.direct_solve(Ao, bo)
    uso = O.direct_solve(*O.poisson_p1_system(p, t, kstar)[:2])
    rhs = Mo @ (uo - uso)
    rhs[d] = 0
    go = O.adjoint_grad_k_p1(p, t, uo, O.direct_solve(Ao, rhs), d)
    assert _rel(grad, go) < 1e-8
    # central FD with defect-correction solves on seeded elements (h = 1e-6 k_e); the loss
    # difference is formed from the increments d+- (L+ - L- = e0^T M (d+ - d-) + (d+Md+ - d-Md-)/2)
    rng = np.random.default_rng(3)
    u0 = u.detach()
    e0 = u0 - ustar
    for el in rng.choice(T, 6, replace=False):
        h = 1e-6 * k[el]
        ds = []
        for sg in (1, -1):
            kk = k.copy()
            kk[el] += sg * h
            A2, b2 = s.evaluate({"k": torch.tensor(kk, device="cuda")})
            ds.append(solve.cg(tol=1e-12)(A2, b2 - A2.mv(u0), M=precond.jacobi()))
        dp, dm = ds
        dL = torch.dot(e0, M.mv(dp - dm)) + 0.5 * (torch.dot(dp, M.mv(dp)) - torch.dot(dm, M.mv(dm)))
        fd = float(dL) / (2 * h)
        assert abs(fd - grad[el]) <= 1e-6 * abs(grad[el]) + 1e-15, (el, fd, grad[el])


class _Exporter:
    """A foreign DLPack exporter (what a JAX array presents): only __dlpack__ and its device."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, *args, **kw):
        return self._t.__dlpack__(*args, **kw)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_linear_slot_over_dlpack(dev):
    """The reference's linear= slot signature with everything passed by DLPack: the operator as
    BCOO triplets with the Dirichlet unit diagonal appended as extra triplets (the reference's
    parametric operator, jno/utils/solver/fem_utils.py:3332-3359), b as an exporter; the solution
    matches the oracle and leaves through DLPack without a copy."""
    from jno_amd import interop, precond, solve

    n = 8
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 41)
    m, s, A, b = _system(n, k)
    data, idx = A.bcoo
    mask = s.dmask.bool()
    rows, cols = idx[:, 0].long(), idx[:, 1].long()
    data2 = data.clone()
    data2[(rows == cols) & mask[rows]] = 0.0  # constrained diagonals zeroed in place ...
    dd = torch.nonzero(mask).flatten().to(torch.int32)
    data3 = torch.cat([data2, torch.ones(dd.numel(), dtype=torch.float64, device="cuda")])
    idx3 = torch.cat([idx, torch.stack([dd, dd], dim=1)])  # ... and the unit diagonal appended
    for precision in ("fp64", "mixed"):
        slot = interop.linear_slot(solve.cg(tol=1e-10, precision=precision))
        x = slot((_Exporter(data3), _Exporter(idx3), A.shape), _Exporter(b), M=precond.jacobi())
        Ao, bo, _ = O.poisson_p1_system(p, t, k)
        assert _rel(_np(x), O.direct_solve(Ao, bo)) <= SOL_TOL, precision
        back = torch.from_dlpack(_Exporter(x))
        assert back.data_ptr() == x.data_ptr()


def test_autotuner_choices(dev):
    """The format choice (DIA on the lattice, CSR with a lane count off it), the launch-shape tuner
    (installs the fastest of the candidates, and the solve still matches the oracle) and the
    batch chunk tuner."""
    from jno_amd import autotune, precond, solve
    from jno_amd.batch import BatchSolver

    n = 9
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 51)
    m, s, A, b = _system(n, k)
    assert autotune.choose_format(A) == ("dia", 1)
    # off the lattice the non-axis couplings are nonzero: more diagonals, or CSR with tuned lanes
    pj = p.copy()
    interior = np.ones(p.shape[0], bool)
    interior[O.boundary_nodes(p, t)] = False
    pj[interior] += np.random.default_rng(5).uniform(-0.2, 0.2, (interior.sum(), 3)) / n
    import jno_amd as J

    mj = J.Mesh(points=torch.tensor(pj, device="cuda"), cells=m.cells)
    Aj, bj = J.fem_poisson(mj).evaluate({"k": torch.tensor(k, device="cuda")})
    fmt, lanes = autotune.choose_format(Aj, allow=("csr",))
    assert fmt == "csr" and lanes in autotune.CANDIDATE_LANES
    shapes = ((256, 64, 0, 5), (512, 128, 3, 4), (1024, 64, 0, 1))
    defaults = solve.get_tuning("mixed")
    best = autotune.tune_launch(A, b, precision="mixed", iters=40, shapes=shapes)
    # the shape is attached to this operator, not installed process-wide
    assert best in shapes and solve.get_tuning("mixed") == defaults
    op = solve.cg(tol=1e-10, precision="mixed").operator(A)
    assert op.tuning is not None and op.tuning.mixed == best
    u = solve.cg(tol=1e-10, precision="mixed")(A, b, M=precond.jacobi())
    Ao, bo, _ = O.poisson_p1_system(p, t, k)
    assert _rel(_np(u), O.direct_solve(Ao, bo)) <= SOL_TOL
    systems = [_system(nn, O.k_lognormal(6 * nn**3, 60 + nn))[2:] for nn in (5, 7, 6)]
    kw = autotune.tune_batch(systems, chunks=(2, 8), shapes=((2048, 8), (4096, 2048)), tol=1e-10)
    assert kw["steps_per_graph"] in (2, 8) and (kw["rows_per_entry"], kw["grid"]) in ((2048, 8), (4096, 2048))
    xs = BatchSolver(tol=1e-10, **kw)(systems)
    for (Ai, bi), xi in zip(systems, xs):
        # the solver stops on its own FP64 residual at 1e-10; recomputed here in another order
        assert float(torch.linalg.norm(bi - Ai.mv(xi)) / torch.linalg.norm(bi)) <= 1.5e-10


@pytest.mark.parametrize("n", [2, 4, 8, 32])
def test_p_eager_bcoo_bytes(dev, n):
    # the reference's stored operator (compress_eager after the Dirichlet edit): the BCOO index array
    # byte for byte, the closed-form entry count, the values to the assembly bar (dyadic spacings:
    # there the vanishing couplings are exact zeros in any rounding)
    import jno_amd as J

    m = J.kuhn_box(n)
    p, t = O.kuhn_box(n)
    k = O.k_sin(p, t)
    A, b = J.fem_poisson(m).assembled({"k": torch.as_tensor(k, device="cuda")})
    Ao, bo, d = O.poisson_p1_system(p, t, k)
    data_o, idx_o = O.bcoo_view(Ao)
    assert isinstance(A, J.BCOO) and A.indices.dtype == torch.int32 and A.indices.shape == (A.nse, 2)
    assert A.nse == idx_o.shape[0] == O.p_eager_nse_kuhn(n)
    assert _np(A.indices).tobytes() == idx_o.tobytes()
    assert np.abs(_np(A.data) - data_o).max() <= VAL_TOL * np.abs(data_o).max()
    assert np.all(_np(A.data) != 0.0)
    if n == 32:
        import json

        meta = json.load(open("tests/golden/pattern_sha256.json"))
        import hashlib

        assert hashlib.sha256(_np(A.indices).astype("<i4").tobytes()).hexdigest() == meta["p1_eager_n32"]["indices_sha256"]


def test_p_eager_bcoo_golden_64_and_nonzero_g(dev):
    # 64^3 against the committed hash; a nonzero g and a nodal load leave the pattern unchanged, and
    # the BCOO solves in the linear= slot like the CSR
    import hashlib
    import json

    import jno_amd as J

    meta = json.load(open("tests/golden/pattern_sha256.json"))
    m = J.kuhn_box(64)
    p = _np(m.points)
    k = torch.as_tensor(O.k_sin(p, _np(m.cells)), device="cuda")
    f = torch.as_tensor(1.0 + p[:, 0], device="cuda")
    A, b = J.fem_poisson(m, f=f, g=0.25).assembled({"k": k})
    assert A.nse == meta["p1_eager_n64"]["nse"]
    assert hashlib.sha256(_np(A.indices).astype("<i4").tobytes()).hexdigest() == meta["p1_eager_n64"]["indices_sha256"]
    x = J.interop.linear_slot(J.solve.cg(tol=1e-10))(A, b, M=J.precond.jacobi())
    A2, b2 = J.fem_poisson(m, f=f, g=0.25).evaluate({"k": k})
    x2 = J.solve.cg(tol=1e-10)(A2, b2, M=J.precond.jacobi())
    assert _rel(_np(x), _np(x2)) < 1e-9


def test_dia_checked_fill_rejects_a_missing_lower_twin(dev):
    # a structurally unsymmetric CSR: the upper entry (0, 1) is stored, its lower twin (1, 0) is not;
    # the DIA layout would read A[1, 0] through symmetry, so the checked fill must refuse it
    import ctypes as C

    from jno_amd import _lib

    n = 4
    rowptr = torch.tensor([0, 2, 3, 5, 7], dtype=torch.int32, device="cuda")
    col = torch.tensor([0, 1, 1, 2, 3, 2, 3], dtype=torch.int32, device="cuda")
    vals = torch.tensor([2.0, -1.0, 2.0, 2.0, -1.0, -1.0, 2.0], dtype=torch.float64, device="cuda")
    ld = 64
    d64 = torch.empty(2 * ld, dtype=torch.float64, device="cuda")
    ws = torch.empty(256, dtype=torch.uint8, device="cuda")
    offs = (C.c_int32 * _lib.JNO_DIA_MAX)(0, 1)
    ok = C.c_int(1)
    _lib.call("jno_op_fill_dia_checked", n, _lib.ptr(rowptr), _lib.ptr(col), _lib.ptr(vals), 2, offs, _lib.ptr(d64),
              None, C.byref(ok), _lib.ptr(ws), 256, _lib.stream_ptr())
    assert ok.value == 0
    # with the twin present (and equal) the same offsets are accepted
    rowptr2 = torch.tensor([0, 2, 4, 6, 8], dtype=torch.int32, device="cuda")
    col2 = torch.tensor([0, 1, 0, 1, 2, 3, 2, 3], dtype=torch.int32, device="cuda")
    vals2 = torch.tensor([2.0, -1.0, -1.0, 2.0, 2.0, -1.0, -1.0, 2.0], dtype=torch.float64, device="cuda")
    _lib.call("jno_op_fill_dia_checked", n, _lib.ptr(rowptr2), _lib.ptr(col2), _lib.ptr(vals2), 2, offs,
              _lib.ptr(d64), None, C.byref(ok), _lib.ptr(ws), 256, _lib.stream_ptr())
    assert ok.value == 1


def test_concurrent_solves_with_per_call_tuning(dev):
    """Two threads solve different systems at the same time, each on its own stream with its own
    launch shape (per-call tuning, no process-wide state): every result equals, bit for bit, the
    same solve run alone, and the library defaults are untouched."""
    import threading

    from jno_amd import precond, solve

    cases = []
    for n, seed, prec, shape in ((12, 21, "mixed", (512, 512, 0, 5)), (9, 22, "mixed", (256, 768, 2, 5)),
                                 (10, 23, "fp64", (1024, 64, 0, 1)), (8, 24, "fp64", (256, 128, 0, 0))):
        m, s, A, b = _system(n, O.k_lognormal(6 * n**3, seed))
        sv = solve.cg(tol=1e-10, precision=prec, tuning=solve.Tuning.defaults().with_shape(prec, *shape))
        cases.append((A, b, sv))
    defaults = (solve.get_tuning("fp64"), solve.get_tuning("mixed"))
    ref = [_np(sv(A, b, M=precond.jacobi())) for A, b, sv in cases]
    out = [[None] * 3 for _ in cases]
    errors = []

    def work(i):
        try:
            A, b, sv = cases[i]
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for r in range(3):
                    out[i][r] = _np(sv(A, b, M=precond.jacobi()))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((i, repr(e)))

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, r in enumerate(ref):
        for x in out[i]:
            assert np.array_equal(x, r), i
    assert (solve.get_tuning("fp64"), solve.get_tuning("mixed")) == defaults


def test_plan_cache_is_keyed_on_the_cells_content(dev):
    # compress_plan's memo (jno/utils/solver/fem_utils.py:4027-4060): the same connectivity gives the
    # same plan, a different one misses; the points do not enter the key (the pattern depends on the
    # cells only), and the fused DIA assembly still checks its layout on the new values
    import ctypes as C

    import jno_amd as J
    from jno_amd import _lib, fem, precond, solve

    fem.clear_plan_cache()
    m1, m2, m3 = J.kuhn_box(6), J.kuhn_box(6), J.kuhn_box(7)
    assert m1.cells is not m2.cells
    assert m1.plan(1) is m2.plan(1) and m1.plan(1) is not m3.plan(1)
    assert m1.plan(3) is not m1.plan(1)
    # content hash and equality
    ws = torch.empty(256, dtype=torch.uint8, device="cuda")
    h1, h2 = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
    c = m1.cells.clone()
    _lib.call("jno_hash_i32", _lib.ptr(c), c.numel(), h1, _lib.ptr(ws), 256, _lib.stream_ptr())
    c.view(-1)[[5, 9]] = c.view(-1)[[9, 5]]  # two entries swapped: same multiset, other content
    _lib.call("jno_hash_i32", _lib.ptr(c), c.numel(), h2, _lib.ptr(ws), 256, _lib.stream_ptr())
    assert tuple(h1) != tuple(h2)
    eq = C.c_int(0)
    _lib.call("jno_equal_i32", _lib.ptr(c), _lib.ptr(m1.cells), c.numel(), C.byref(eq), _lib.ptr(ws), 256,
              _lib.stream_ptr())
    assert eq.value == 0
    # a perturbed mesh on the same cells shares the plan and still solves to the oracle
    n = 6
    p, t = O.kuhn_box(n)
    pj = p.copy()
    interior = np.ones(p.shape[0], bool)
    interior[O.boundary_nodes(p, t)] = False
    pj[interior] += np.random.default_rng(9).uniform(-0.2, 0.2, (interior.sum(), 3)) / n
    k = O.k_lognormal(t.shape[0], 13)
    sv = solve.cg(tol=1e-10, precision="mixed")
    A1, b1 = J.fem_poisson(m1).evaluate({"k": torch.tensor(k, device="cuda")})
    sv(A1, b1, M=precond.jacobi())  # the DIA offsets of this plan are learnt here
    mj = J.Mesh(points=torch.tensor(pj, device="cuda"), cells=m2.cells.clone())
    assert mj.plan(1) is m1.plan(1)
    Aj, bj = J.fem_poisson(mj).evaluate({"k": torch.tensor(k, device="cuda")})
    Ao, bo, _ = O.poisson_p1_system(pj, t, k)
    assert _rel(_np(sv(Aj, bj, M=precond.jacobi())), O.direct_solve(Ao, bo)) <= SOL_TOL


@pytest.mark.parametrize("fmt", ["dia", "csr"])
def test_mixed_solve_with_a_zero_diagonal_row(dev, fmt):
    """A row and column of zeros (b = 0 there): the Jacobi guard leaves the row unscaled
    (precond.py:107, 1/d only for |d| > 1e-30), so the Jacobi-scaled FP32 iteration does not apply
    (its scaled main diagonal would not be 1) and the solve falls back to the unscaled one, which
    the device finds before the first iteration.  The solution equals the oracle's on the remaining
    rows, 0 on the zero row, in both precisions."""
    from jno_amd import precond, solve
    from jno_amd.fem import CSR

    n = 7
    p, t = O.kuhn_box(n)
    k = O.k_lognormal(t.shape[0], 41)
    m, s, A, b = _system(n, k)
    N = A.shape[0]
    j = N // 2 + 3
    rp, col, v = _np(A.rowptr), _np(A.col), _np(A.vals).copy()
    rows = np.repeat(np.arange(N), np.diff(rp))
    v[(rows == j) | (col == j)] = 0.0
    bj = _np(b).copy()
    bj[j] = 0.0
    Az = CSR(A.rowptr, A.col, torch.tensor(v, device="cuda"), A.shape)
    bz = torch.tensor(bj, device="cuda")
    keep = np.arange(N) != j
    Ao = O.csr_matrix(rp, col, v, N)[keep][:, keep]
    uo = np.zeros(N)
    uo[keep] = O.direct_solve(Ao.tocsr(), bj[keep])
    for prec in ("mixed", "fp64"):
        sv = solve.cg(tol=1e-10, precision=prec, format=fmt)
        u = _np(sv(Az, bz, M=precond.jacobi()))
        assert sv.last_stats["relres"] <= 1e-10, (prec, sv.last_stats)
        assert u[j] == 0.0
        assert _rel(u, uo) <= SOL_TOL, (prec, fmt)
