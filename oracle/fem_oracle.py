"""CPU restatement of jNO's Lagrange P1/P2 finite-element hot path (NumPy/SciPy, FP64).

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product path (``jno_amd``) never
calls it and must fail loudly when its HIP library is missing.

Every function restates a piece of the reference (FhG-IISB/jNO 0.3.1) and cites the file:line it
follows.  Paths are relative to the reference checkout.

Pinning (what checks this restatement beyond itself; see tests/test_oracle_*.py):
  * kuhn_box              closed-form counts, orientation-free volume sum, facet census
  * promote_p2            loop restatement == vectorised form; (2n+1)^3 node count; midpoint identity
  * pattern_struct        closed-form nnz per config; brute-force dense adjacency on tiny meshes
  * p1 element / load     linear-field exactness (reference tests/test_fem_3d.py:84-97), row sums 0,
                          symmetry, one-hot element support (tests/test_topology_optimisation.py:210-237)
  * mass_p1               closed form integral check: 1^T M 1 == volume
  * elasticity_p2         rigid-body null space, symmetry, quadratic-field exactness
                          (tests/test_fem_higher_order.py:133-150 restated), P1-on-linear agreement
  * dirichlet_eliminate   identities of symmetric elimination (tests/test_fem_dirichlet_symmetry.py)
  * bcoo_view             closed-form nse on Kuhn lattices; brute-force dense nonzero positions
  * theta_* (transient)   amplification factor on discrete eigenmodes, steady limit, analytic decay,
                          manufactured forced solutions; theta_march_grad_k by central FD
  * pcg                   scipy.sparse.linalg.spsolve (direct, independent of any iteration)
  * adjoint_grad_k        central finite differences with defect-correction solves
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------------------------
# A1  Kuhn lattice mesher
# ---------------------------------------------------------------------------------------------

#: Six Kuhn/Freudenthal tets per voxel sharing the 000->111 diagonal; corner c encodes
#: (di, dj, dk) = (c>>2 & 1, c>>1 & 1, c & 1).  jno/domain/geometries.py:245
KUHN = ((0, 4, 6, 7), (0, 4, 5, 7), (0, 2, 6, 7), (0, 2, 3, 7), (0, 1, 5, 7), (0, 1, 3, 7))

FACE_NAMES = ("left", "right", "bottom", "top", "front", "back")


def kuhn_box(nx, ny=None, nz=None, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Points (N,3) f64 and tets (T,4) int64 of the structured box.

    Follows ``Geometries.equi_distant_box`` (jno/domain/geometries.py:204-301): per-axis
    ``linspace``, ``meshgrid(indexing="ij")``, node id ``(i*(ny+1)+j)*(nz+1)+k``, voxels in (i,j,k)
    loop order with the six Kuhn tets of each voxel in ``KUHN`` order.  Vectorised, same output.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    X = np.linspace(lo[0], hi[0], nx + 1)
    Y = np.linspace(lo[1], hi[1], ny + 1)
    Z = np.linspace(lo[2], hi[2], nz + 1)
    xx, yy, zz = np.meshgrid(X, Y, Z, indexing="ij")
    points = np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])
    ny1, nz1 = ny + 1, nz + 1
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    corner = np.stack(
        [((i + ((c >> 2) & 1)) * ny1 + (j + ((c >> 1) & 1))) * nz1 + (k + (c & 1)) for c in range(8)], axis=1
    )
    tets = np.stack([corner[:, list(t)] for t in KUHN], axis=1).reshape(-1, 4).astype(np.int64)
    return points, tets


def kuhn_box_faces(nx, ny=None, nz=None):
    """Boundary triangles per named face, as in jno/domain/geometries.py:262-285 (quad -> 2 tris)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    ny1, nz1 = ny + 1, nz + 1

    def idx(i, j, k):
        return (i * ny1 + j) * nz1 + k

    def quad_tris(node, na, nb):
        a, b = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
        a, b = a.ravel(), b.ravel()
        p00, p10, p11, p01 = node(a, b), node(a + 1, b), node(a + 1, b + 1), node(a, b + 1)
        t = np.stack([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)], axis=1)
        return t.reshape(-1, 3).astype(np.int64)

    return {
        "left": quad_tris(lambda j, k: idx(0, j, k), ny, nz),
        "right": quad_tris(lambda j, k: idx(nx, j, k), ny, nz),
        "bottom": quad_tris(lambda i, k: idx(i, 0, k), nx, nz),
        "top": quad_tris(lambda i, k: idx(i, ny, k), nx, nz),
        "front": quad_tris(lambda i, j: idx(i, j, 0), nx, ny),
        "back": quad_tris(lambda i, j: idx(i, j, nz), nx, ny),
    }


# ---------------------------------------------------------------------------------------------
# A2  P2 promotion
# ---------------------------------------------------------------------------------------------

#: basix P2-tet interpolation points: the 4 vertices, then the midpoints of edges
#: (2,3),(1,3),(1,2),(0,3),(0,2),(0,1) (jno/utils/solver/fem_topology.py:44).
P2_EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
P2_REF = np.array(
    [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    + [[0.5 * (a == 1 or b == 1), 0.5 * (a == 2 or b == 2), 0.5 * (a == 3 or b == 3)] for a, b in P2_EDGES],
    dtype=float,
)


def _p2_phys(points, cells):
    w = np.empty((10, 4))
    w[:, 0] = 1.0 - P2_REF.sum(axis=1)
    w[:, 1:] = P2_REF
    return np.einsum("dc,ncg->ndg", w, points[cells])


def promote_p2_loop(points, cells):
    """Literal restatement of ``_promote_to_degree`` (jno/utils/solver/fem_utils.py:1979-2044):
    coordinate hash ``round(x/tol)`` with ``tol = 1e-7*extent``, vertices seeded first, new ids in
    cell-major / local-DOF-minor first-appearance order.  Pure Python: small meshes only."""
    points = np.asarray(points, float)
    phys = _p2_phys(points, cells)
    extent = float(np.max(points.max(axis=0) - points.min(axis=0)))
    tol = 1e-7 * (extent or 1.0)
    key = lambda p: tuple(np.round(np.asarray(p) / tol).astype(np.int64))  # noqa: E731
    gid = {}
    out = []
    for v in range(points.shape[0]):
        gid[key(points[v])] = v
        out.append(points[v])
    cells_k = np.zeros((cells.shape[0], 10), dtype=np.int64)
    nid = points.shape[0]
    for c in range(cells.shape[0]):
        for d in range(10):
            kk = key(phys[c, d])
            g = gid.get(kk)
            if g is None:
                g = gid[kk] = nid
                out.append(phys[c, d])
                nid += 1
            cells_k[c, d] = g
    return np.asarray(out), cells_k


def promote_p2(points, cells):
    """Vectorised form of :func:`promote_p2_loop` (same ids, checked equal by the tests)."""
    points = np.asarray(points, float)
    phys = _p2_phys(points, cells).reshape(-1, 3)
    extent = float(np.max(points.max(axis=0) - points.min(axis=0)))
    tol = 1e-7 * (extent or 1.0)
    keys = np.round(np.concatenate([points, phys]) / tol).astype(np.int64)
    uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    nv = points.shape[0]
    # a unique key seen first at position < nv is an original vertex and keeps its id
    order = np.argsort(first, kind="stable")
    new_id = np.empty(uniq.shape[0], dtype=np.int64)
    is_vert = first[order] < nv
    new_id[order[is_vert]] = first[order[is_vert]]
    n_new = int((~is_vert).sum())
    new_id[order[~is_vert]] = nv + np.arange(n_new)
    allpts = np.concatenate([points, phys])
    out = np.empty((nv + n_new, 3))
    out[new_id] = allpts[first]
    cells_k = new_id[inv[nv:]].reshape(-1, 10)
    return out, cells_k


# ---------------------------------------------------------------------------------------------
# A3, A7  DOF map and the structural CSR pattern
# ---------------------------------------------------------------------------------------------


def cell_dofs(cells, vec=1):
    """dof = node*vec + comp, local order node-major then comp (jno/utils/solver/fem_native.py:907-912)."""
    cells = np.asarray(cells, np.int64)
    return (cells[:, :, None] * vec + np.arange(vec)).reshape(cells.shape[0], -1)


def pattern_struct(cells, vec=1, ndof=None):
    """Sorted unique (row, col) of every element pair = the CSR index structure.

    Restates the raw append order of jno/utils/solver/fem_native.py:2264-2298 (rows
    ``cdofs[:, :, None]``, cols ``cell_all_dofs[:, None, :]``) and ``compress_plan``
    (jno/utils/solver/fem_utils.py:4009-4060): key ``row*(max_col+1)+col``, ``np.unique`` with the
    inverse map.  Returns (rowptr int64, col int32, inverse int64, nse)."""
    cd = cell_dofs(cells, vec)
    T, m = cd.shape
    rows = np.broadcast_to(cd[:, :, None], (T, m, m)).reshape(-1)
    cols = np.broadcast_to(cd[:, None, :], (T, m, m)).reshape(-1)
    stride = int(cols.max()) + 1
    uniq, inv = np.unique(rows * stride + cols, return_inverse=True)
    r = uniq // stride
    c = (uniq % stride).astype(np.int32)
    if ndof is None:
        ndof = int(cd.max()) + 1
    rowptr = np.zeros(ndof + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return rowptr, c, inv.reshape(-1), int(uniq.shape[0])


def csr_matrix(rowptr, col, vals, n):
    return sp.csr_matrix((np.asarray(vals, float), np.asarray(col), np.asarray(rowptr)), shape=(n, n))


# ---------------------------------------------------------------------------------------------
# A4-A6, A9, A19  element values
# ---------------------------------------------------------------------------------------------

_GREF_P1 = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def p1_geometry(points, cells):
    """Per-cell physical gradients G (T,4,3) and volume |detJ|/6.

    J has columns v1-v0, v2-v0, v3-v0; grad(phi) = dphi/dxi . inv(J); measure abs(detJ)
    (``_cell_fields``, jno/utils/solver/fem_native.py:1461-1497; ``identity_pushforward``,
    jno/utils/solver/fem_lagrange.py:348-381)."""
    v = points[cells[:, :4]]
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)
    det = np.linalg.det(J)
    K = np.linalg.inv(J)
    G = np.einsum("ak,tkj->taj", _GREF_P1, K)
    return G, np.abs(det) / 6.0


def p1_stiffness(points, cells, k):
    """K_e = k_e |T_e| G G^T -- exact value of ``jacfwd(_vol_elem_res)`` for the P1 form
    k grad u . grad v (jno/utils/solver/fem_native.py:1645-1691, :2331-2345)."""
    G, vol = p1_geometry(points, cells)
    return np.einsum("t,taj,tbj->tab", np.asarray(k, float) * vol, G, G)


def p1_mass(points, cells):
    """Consistent P1 mass |T|/20 (1 + delta_ij): the u*v form, exact under the degree-2 rule."""
    _, vol = p1_geometry(points, cells)
    return vol[:, None, None] * (np.ones((4, 4)) + np.eye(4))[None] / 20.0


def p1_load(points, cells, f, n):
    """b_i = sum_e int f phi_i (``R.at[cdofs].add``, jno/utils/solver/fem_native.py:2145-2179).
    ``f`` is a scalar (f*|T|/4 per vertex) or a nodal P1 field (b = M f, exact)."""
    _, vol = p1_geometry(points, cells)
    f = np.asarray(f, float)
    if f.ndim == 0:
        contrib = np.repeat(vol[:, None] * float(f) / 4.0, 4, axis=1)
    else:
        fe = f[cells]
        contrib = vol[:, None] * (fe + fe.sum(axis=1, keepdims=True)) / 20.0
    return np.bincount(cells.reshape(-1), weights=contrib.reshape(-1), minlength=n)


def assemble(elem_mats, cells, vec, ndof, plan=None):
    """segment_sum of raw element triplets into the plan's unique slots
    (jno/utils/solver/fem_native.py:2318-2329; ``apply_compress_plan``, fem_utils.py:4094-4101)."""
    if plan is None:
        plan = pattern_struct(cells, vec, ndof)
    rowptr, col, inv, nse = plan
    vals = np.bincount(inv, weights=np.asarray(elem_mats).reshape(-1), minlength=nse)
    return rowptr, col, vals


# --- P2 Lagrange on the reference tet (basix DOF order) -------------------------------------

#: Degree-2 exact 4-point rule on the reference tet (Keast/Hammer), weights sum to 1/6.  The
#: reference uses basix's default degree-2 rule; any degree-2-exact rule gives the same element
#: matrices up to rounding because every integrand here has degree <= 2.
_QA, _QB = 0.1381966011250105, 0.5854101966249685
QUAD_P2 = np.array([[_QA, _QA, _QA], [_QB, _QA, _QA], [_QA, _QB, _QA], [_QA, _QA, _QB]])
QUAD_W = np.full(4, 1.0 / 24.0)


def p2_ref_grads(xi):
    """d phi / d xi (10,3) of the P2 basis at reference point xi, basix DOF order."""
    l = np.array([1.0 - xi.sum(), xi[0], xi[1], xi[2]])
    dl = np.array([[-1.0, -1.0, -1.0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    g = np.empty((10, 3))
    for a in range(4):
        g[a] = (4.0 * l[a] - 1.0) * dl[a]
    for e, (a, b) in enumerate(P2_EDGES):
        g[4 + e] = 4.0 * (l[a] * dl[b] + l[b] * dl[a])
    return g


def p2_ref_vals(xi):
    l = np.array([1.0 - xi.sum(), xi[0], xi[1], xi[2]])
    v = np.empty(10)
    v[:4] = l * (2 * l - 1)
    for e, (a, b) in enumerate(P2_EDGES):
        v[4 + e] = 4 * l[a] * l[b]
    return v


def lame(nu):
    """Lame parameters for E=1: lambda = nu/((1+nu)(1-2nu)), mu = 1/(2(1+nu))."""
    return nu / ((1 + nu) * (1 - 2 * nu)), 1.0 / (2 * (1 + nu))


def elasticity_p2(points, cells10, E, nu):
    """(T,30,30) K_e = int E [lam d_i phi_a d_j phi_b + mu (d_i phi_b d_j phi_a + delta_ij grad phi_a.grad phi_b)]
    for the form E (lam tr eps(u) tr eps(v) + 2 mu eps(u):eps(v))
    (tests/test_topology_optimisation.py:59-66 weak form; element Jacobian as in
    jno/utils/solver/fem_native.py:2331-2345), local dof = node*3 + comp."""
    lam, mu = lame(nu)
    v = points[cells10[:, :4]]
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)
    det = np.abs(np.linalg.det(J))
    Kinv = np.linalg.inv(J)
    T = cells10.shape[0]
    Ke = np.zeros((T, 10, 3, 10, 3))
    I3 = np.eye(3)
    for q in range(4):
        gr = p2_ref_grads(QUAD_P2[q])
        G = np.einsum("ak,tkj->taj", gr, Kinv)  # (T,10,3)
        w = QUAD_W[q] * det * np.asarray(E, float)
        Ke += w[:, None, None, None, None] * (
            lam * np.einsum("tai,tbj->taibj", G, G)
            + mu * np.einsum("tbi,taj->taibj", G, G)
            + mu * np.einsum("tak,tbk,ij->taibj", G, G, I3)
        )
    return Ke.reshape(T, 30, 30)


def facet_traction_p2(points, tris6, t):
    """Load of a constant traction t on P2 boundary triangles: int_F t phi_i dS.
    For a straight 6-node triangle int phi_vertex = 0 and int phi_edge = |F|/3
    (surface term, jno/utils/solver/fem_native.py:1882-1960).  tris6 = (F,6): 3 vertices then
    the 3 edge midpoints.  Returns (F, 6, 3) contributions."""
    p = points[tris6[:, :3]]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    w = np.zeros((tris6.shape[0], 6))
    w[:, 3:] = area[:, None] / 3.0
    return w[:, :, None] * np.asarray(t, float)[None, None, :]


# ---------------------------------------------------------------------------------------------
# A10, A11  boundary facets, Dirichlet sets, symmetric elimination
# ---------------------------------------------------------------------------------------------

_TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def _facet_edge_dofs(face):
    """P2 local dofs on reference facet ``face``: the edges among its three vertices."""
    return [4 + e for e, (a, b) in enumerate(P2_EDGES) if a in face and b in face]


def boundary_facets(cells):
    """Facets used by exactly one cell, as global node rows (vertices first, then P2 facet nodes).
    ``_boundary_facets`` (jno/_fem.py:2695-2773): combo-major facet list, canonical sort,
    ``np.unique(return_index, return_counts)``, keep count == 1."""
    cells = np.asarray(cells, np.int64)
    T = cells.shape[0]
    allf = np.concatenate([cells[:, list(c)] for c in _TET_FACES], axis=0)
    canon = np.sort(allf, axis=1)
    _u, idx, counts = np.unique(canon, axis=0, return_index=True, return_counts=True)
    bidx = idx[counts == 1]
    if cells.shape[1] == 4:
        return allf[bidx]
    table = np.array([list(c) + _facet_edge_dofs(c) for c in _TET_FACES], dtype=np.int64)
    return cells[(bidx % T)[:, None], table[bidx // T]]


def boundary_nodes(points, cells, region="boundary", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Ascending node ids on the boundary (``_boundary_node_ids``,
    jno/utils/solver/fem_native.py:2563-2636): every node of a boundary facet; a named box face
    keeps those whose coordinate lies on that face."""
    bn = np.unique(boundary_facets(cells).reshape(-1))
    if region == "boundary":
        return bn
    ax = {"left": (0, lo[0]), "right": (0, hi[0]), "bottom": (1, lo[1]), "top": (1, hi[1]),
          "front": (2, lo[2]), "back": (2, hi[2])}[region]
    return bn[points[bn, ax[0]] == ax[1]]


def dirichlet_pairs(nodes, vec=1, comps=None, g=0.0):
    """(dof, value) pairs: dof = node*vec + comp (jno/utils/solver/fem_native.py:2638-2711)."""
    comps = range(vec) if comps is None else comps
    dofs = np.array([n * vec + c for n in nodes for c in comps], dtype=np.int64)
    vals = np.broadcast_to(np.asarray(g, float), dofs.shape).copy()
    return dofs, vals


def dirichlet_eliminate(A, b, dofs, vals):
    """Symmetric elimination (``_apply_dirichlet_symmetric``, jno/utils/solver/fem_1d.py:667-686 with
    ``bcoo_zero_rows_cols`` + ``bcoo_set_unit_diag``, jno/utils/solver/fem_utils.py:3332-3359):
    b <- b - A e;  data <- data (1-m_row)(1-m_col);  A[d,d] <- 0 + 1;  b[d] <- g.
    The pattern is unchanged: constrained couplings become stored zeros."""
    A = A.tocsr(copy=True)
    A.sort_indices()
    n = A.shape[0]
    e = np.zeros(n)
    e[dofs] = vals
    b = np.asarray(b, float) - A @ e
    isd = np.zeros(n)
    isd[dofs] = 1.0
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    keep = (1.0 - isd[rows]) * (1.0 - isd[A.indices])
    A.data = A.data * keep
    on_d = (rows == A.indices) & (isd[rows] > 0)
    A.data[on_d] += 1.0
    b[dofs] = vals
    return A, b


def compress_eager(A):
    """Drop entries whose summed value is exactly 0.0 (``compress_eager``,
    jno/utils/solver/fem_utils.py:4063-4091); returns a sorted CSR (P_eager)."""
    A = A.tocsr(copy=True)
    A.sort_indices()
    keep = np.abs(A.data) > 0.0
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    return sp.csr_matrix((A.data[keep], (rows[keep], A.indices[keep])), shape=A.shape)


def bcoo_view(A):
    """(data, indices) of the BCOO that ``compress_eager`` returns for the assembled operator
    (jno/utils/solver/fem_utils.py:4063-4091 after ``compress_plan``'s ``np.unique`` of the flattened
    key row * stride + col, fem_utils.py:4049-4053): one entry per (row, col), sorted row-major,
    exact zeros dropped, indices int32 (nse, 2)."""
    A = compress_eager(A)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    idx = np.stack([rows, A.indices], axis=1).astype(np.int32)
    return A.data.copy(), idx


def p_eager_nse_kuhn(n):
    """Closed-form nse of the eliminated P1 Poisson operator on the n^3 Kuhn lattice with the whole
    boundary constrained: the Kuhn P1 stiffness couples only axis neighbours (the face and body
    diagonals are exact zeros), elimination leaves one diagonal entry per boundary row, and each
    interior row keeps its diagonal and its interior axis neighbours."""
    nb = (n + 1) ** 3 - (n - 1) ** 3
    return nb + (n - 1) ** 3 + 6 * (n - 1) ** 2 * (n - 2)


# ---------------------------------------------------------------------------------------------
# A14-A17  Jacobi, PCG with JAX semantics, residual guard
# ---------------------------------------------------------------------------------------------


def jacobi_inv(A):
    """inv = 1/d with the guard |d| > 1e-30 else 1 (jno/precond.py:105-109, jno/utils/solver/linear.py:88-108)."""
    d = A.diagonal()
    return 1.0 / np.where(np.abs(d) > 1e-30, d, 1.0)


def pcg(A, b, dinv=None, tol=1e-10, atol=0.0, maxiter=20000, x0=None):
    """Preconditioned CG with the semantics of ``jax.scipy.sparse.linalg.cg`` as called by
    ``jno.solve.cg`` (jno/solve.py:191-206): x0 = 0, r = b - A x0, z = M r, p = z, gamma = r.z;
    loop while r.r > max(tol^2 b.b, atol^2) and k < maxiter.  Returns (x, iterations)."""
    b = np.asarray(b, float)
    dinv = np.ones_like(b) if dinv is None else dinv
    x = np.zeros_like(b) if x0 is None else np.array(x0, float)
    r = b - A @ x
    z = dinv * r
    p = z.copy()
    gamma = r @ z
    atol2 = max(tol * tol * (b @ b), atol * atol)
    k = 0
    while (r @ r) > atol2 and k < maxiter:
        Ap = A @ p
        alpha = gamma / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = dinv * r
        gamma_new = r @ z
        p = z + (gamma_new / gamma) * p
        gamma = gamma_new
        k += 1
    return x, k


def residual_check(A, b, x, rtol=1e-4):
    """``_maybe_residual_check`` (jno/utils/solver/solver_api.py:224-240)."""
    rel = float(np.linalg.norm(b - A @ x) / (np.linalg.norm(b) + 1e-30))
    if not np.isfinite(rel) or rel > rtol:
        raise RuntimeError(f"jno.solve.cg did not solve the system (relative residual {rel:.1e})")
    return rel


def direct_solve(A, b):
    return spla.spsolve(A.tocsc(), b)


# ---------------------------------------------------------------------------------------------
# Inputs of the benchmark configurations
# ---------------------------------------------------------------------------------------------


def centroids(points, cells):
    return points[cells[:, :4]].mean(axis=1)


def k_sin(points, cells):
    """k_e = 1 + 0.5 sin(2 pi x_c) sin(2 pi y_c) sin(2 pi z_c) at the element centroid."""
    c = centroids(points, cells)
    return 1.0 + 0.5 * np.sin(2 * np.pi * c[:, 0]) * np.sin(2 * np.pi * c[:, 1]) * np.sin(2 * np.pi * c[:, 2])


def k_lognormal(T, seed, sigma=0.5):
    """k_e = exp(sigma z_e), z ~ N(0,1) iid from numpy default_rng(seed) (PCG64)."""
    return np.exp(sigma * np.random.default_rng(seed).standard_normal(T))


# ---------------------------------------------------------------------------------------------
# P1 Poisson system and its adjoint (A13, A18, A19)
# ---------------------------------------------------------------------------------------------


def poisson_p1_system(points, cells, k, f=1.0, region="boundary", g=0.0, plan=None):
    """``FemLinearSystem.evaluate({"k": k})`` for -div(k grad u) = f with u = g on ``region``
    (jno/trace/__init__.py:3612-3615 -> jno/utils/solver/fem_native.py:3311-3318): assembled,
    Dirichlet-eliminated CSR A (P_struct pattern) and b."""
    n = points.shape[0]
    rowptr, col, vals = assemble(p1_stiffness(points, cells, k), cells, 1, n, plan)
    A = csr_matrix(rowptr, col, vals, n)
    b = p1_load(points, cells, f, n)
    dofs, dv = dirichlet_pairs(boundary_nodes(points, cells, region), 1, None, g)
    A, b = dirichlet_eliminate(A, b, dofs, dv)
    return A, b, dofs


def adjoint_grad_k_p1(points, cells, u, lam, dofs):
    """dL/dk_e = -(P lam)_e^T (K_e/k_e) u_e = -|T_e| (G_e P lam_e).(G_e u_e), P zeroing Dirichlet
    dofs: the VJP chain of custom_linear_solve(symmetric=True) -> BCOO data -> Dirichlet mask ->
    segment_sum -> jacfwd (jno/utils/solver/fem_native.py:2326-2345, fem_utils.py:3332-3336)."""
    G, vol = p1_geometry(points, cells)
    lam = np.array(lam, float)
    lam[dofs] = 0.0
    gl = np.einsum("taj,ta->tj", G, lam[cells])
    gu = np.einsum("taj,ta->tj", G, np.asarray(u, float)[cells])
    return -vol * np.einsum("tj,tj->t", gl, gu)


def mass_matrix_p1(points, cells, plan=None):
    n = points.shape[0]
    rowptr, col, vals = assemble(p1_mass(points, cells), cells, 1, n, plan)
    return csr_matrix(rowptr, col, vals, n)


# ---------------------------------------------------------------------------------------------
# P2 vector elasticity system (the 10-node, 3-dof-per-node configuration) and its compliance
# gradient
# ---------------------------------------------------------------------------------------------


def plane_facets_p2(points, cells10, axis, value):
    """P2 boundary facets (F,6) whose three vertices lie on the plane x[axis] == value."""
    bf = boundary_facets(cells10)
    on = np.all(points[bf[:, :3], axis] == value, axis=1)
    return bf[on]


def elasticity_p2_system(points, cells10, E, nu=0.3, traction=(0.0, 0.0, -1.0), clamp=(0, 0.0), load=(0, 1.0),
                         plan=None):
    """A u = b for E (lam tr eps tr eps + 2 mu eps:eps) with u = 0 on the clamp plane (all three
    components) and the constant traction on the load plane (jno/utils/solver/fem_native.py:1882-1960
    surface term; weak form of tests/test_topology_optimisation.py:59-66 with E per element)."""
    n = points.shape[0] * 3
    rowptr, col, vals = assemble(elasticity_p2(points, cells10, E, nu), cells10, 3, n, plan)
    A = csr_matrix(rowptr, col, vals, n)
    b = np.zeros(n)
    tris = plane_facets_p2(points, cells10, *load)
    contrib = facet_traction_p2(points, tris, traction)
    for c in range(3):
        np.add.at(b, tris.reshape(-1) * 3 + c, contrib[:, :, c].reshape(-1))
    nodes = boundary_nodes(points, cells10)
    nodes = nodes[points[nodes, clamp[0]] == clamp[1]]
    dofs, dv = dirichlet_pairs(nodes, 3)
    A, b = dirichlet_eliminate(A, b, dofs, dv)
    return A, b, dofs


def compliance_grad_p2(points, cells10, E, nu, u):
    """dC/dE_e = -u_e^T (K_e/E_e) u_e for the compliance C = b.u (self-adjoint: lam = u)."""
    Ke = elasticity_p2(points, cells10, np.ones(cells10.shape[0]), nu)
    ue = np.asarray(u, float)[cell_dofs(cells10, 3)]
    return -np.einsum("ti,tij,tj->t", ue, Ke, ue)


# ---------------------------------------------------------------------------------------------
# Theta-method transient step of the linear block M u' + A u = c + f(t)
# ---------------------------------------------------------------------------------------------


def transient_dirichlet(M, dofs):
    """The mass half of ``_apply_dirichlet_transient`` (jno/utils/solver/fem_1d.py:939-949):
    the constrained rows AND columns of M are zeroed (no diagonal), the pattern kept.  A and c take
    the steady symmetric elimination (``dirichlet_eliminate``)."""
    M = M.tocsr(copy=True)
    M.sort_indices()
    isd = np.zeros(M.shape[0])
    isd[np.asarray(dofs, np.int64)] = 1.0
    rows = np.repeat(np.arange(M.shape[0]), np.diff(M.indptr))
    M.data = M.data * (1.0 - isd[rows]) * (1.0 - isd[M.indices])
    return M


def theta_operators(Mz, A, dt, theta):
    """Step operator K = Mz + theta dt A (``_add_step_operator``, jno/utils/solver/solver_api.py:
    856-879) and explicit operator E = Mz - (1-theta) dt A, value by value on the shared pattern."""
    Mz = Mz.tocsr()
    A = A.tocsr()
    Mz.sort_indices()
    A.sort_indices()
    if not (np.array_equal(Mz.indptr, A.indptr) and np.array_equal(Mz.indices, A.indices)):
        raise ValueError("M and A must share one pattern")
    K = sp.csr_matrix((Mz.data + (theta * dt) * A.data, A.indices, A.indptr), shape=A.shape)
    E = sp.csr_matrix((Mz.data + (-(1.0 - theta) * dt) * A.data, A.indices, A.indptr), shape=A.shape)
    return K, E


def theta_rhs(Mz, A, c, u, dt, theta, f_now=None, f_next=None):
    """rhs = M u - (1-theta) dt (A u) + dt c + dt (theta f(t+dt) + (1-theta) f(t))
    (``SemidiscreteTimeBlock.step``, jno/utils/solver/backend_blocks.py:342-344)."""
    rhs = Mz @ u - (1.0 - theta) * dt * (A @ u) + dt * np.asarray(c, float)
    if f_next is not None:
        rhs = rhs + dt * (theta * np.asarray(f_next, float) + (1.0 - theta) * np.asarray(f_now, float))
    return rhs


def theta_march(Mz, A, c, u0, dt, theta, n_steps, forcing=None, t0=0.0):
    """States u_0..u_n of ``n_steps`` theta-steps, each solved directly (the exact step answer that
    the reference's per-step Krylov solve at tol 1e-10 approximates, backend_blocks.py:329-351).
    ``forcing(t)`` -> load increment vector, or None."""
    K, _ = theta_operators(Mz, A, dt, theta)
    lu = spla.splu(K.tocsc())
    us = [np.array(u0, float)]
    t = t0
    for _ in range(n_steps):
        fn = (forcing(t), forcing(t + dt)) if forcing is not None else (None, None)
        us.append(lu.solve(theta_rhs(Mz, A, c, us[-1], dt, theta, *fn)))
        t += dt
    return us


def theta_march_tv(Mz, system_at, u0, dt, theta, n_steps, forcing=None, t0=0.0):
    """States u_0..u_n of a march whose operator depends on time: step n uses ``system_at(t_{n+1})``
    = (A, c) for both halves of the theta-step, as ``SemidiscreteTimeBlock.step`` evaluates
    ``operator_fn(t_next, args)`` once per step (jno/utils/solver/backend_blocks.py:335, :342-344;
    the per-step operator of compose_transient_step_solvers, jno/utils/solver/solver_api.py:
    923-935), each step solved directly.  A constant ``system_at`` reproduces ``theta_march``."""
    us = [np.array(u0, float)]
    t = t0
    for _ in range(n_steps):
        A, c = system_at(t + dt)
        K, _ = theta_operators(Mz, A, dt, theta)
        fn = (forcing(t), forcing(t + dt)) if forcing is not None else (None, None)
        us.append(spla.spsolve(K.tocsc(), theta_rhs(Mz, A, c, us[-1], dt, theta, *fn)))
        t += dt
    return us


def theta_march_tv_grad_k(points, cells, Mz, system_at, profile, us, dt, theta, dofs, g, lam_final, t0=0.0):
    """dL/dk_e along a march whose coefficient at time t is profile(t) * k (``system_at(t)`` = the
    eliminated (A, c) at that coefficient): the reverse sweep of ``theta_march_grad_k`` with the
    step operators of each step and the contraction scaled by profile(t_{n+1}), the factor by
    which k_e enters that step's operator.  Returns (dL/dk, dL/du_0)."""
    dofs = np.asarray(dofs, np.int64)
    grad = np.zeros(cells.shape[0])
    lam = np.array(lam_final, float)
    for n in range(len(us) - 2, -1, -1):
        tn1 = t0 + (n + 1) * dt
        A, _ = system_at(tn1)
        K, E = theta_operators(Mz, A, dt, theta)
        mu = spla.spsolve(K.tocsc(), lam)
        w = theta * us[n + 1] + (1.0 - theta) * us[n]
        w[dofs] = g
        grad += dt * profile(tn1) * adjoint_grad_k_p1(points, cells, w, mu, dofs)
        lam = E @ mu
    return grad, lam


def theta_march_grad_k(points, cells, Mz, A, us, dt, theta, dofs, g, lam_final):
    """dL/dk_e for L with dL/du_N = ``lam_final`` along the march ``us``: the reverse sweep
    mu = K^-1 lam_{n+1}, dL/dk += dt * adjoint_grad_k(w_n, mu) with w_n = theta u_{n+1} +
    (1-theta) u_n on free dofs and g on constrained ones, lam_n = E mu (K, E symmetric).  Returns
    (dL/dk, dL/du_0).  The reference gets the same numbers from jax.lax.scan's reverse mode
    (backend_blocks.py:475-530)."""
    K, E = theta_operators(Mz, A, dt, theta)
    lu = spla.splu(K.tocsc())
    dofs = np.asarray(dofs, np.int64)
    grad = np.zeros(cells.shape[0])
    lam = np.array(lam_final, float)
    for n in range(len(us) - 2, -1, -1):
        mu = lu.solve(lam)
        w = theta * us[n + 1] + (1.0 - theta) * us[n]
        w[dofs] = g
        grad += dt * adjoint_grad_k_p1(points, cells, w, mu, dofs)
        lam = E @ mu
    return grad, lam
