"""Assembly front-end: the CSR plan, the assembled operator and ``FemLinearSystem.evaluate``.

Numeric analogue of the reference's parametric FEM operator: a mesh, a per-element coefficient
(a P0 ``jno.np.parameter``), a load and a Dirichlet set go in; the assembled, symmetrically
eliminated CSR ``A`` and load ``b`` come out (``FemLinearSystem.evaluate``,
jno/trace/__init__.py:3612-3615 -> ``_assemble_at``, jno/utils/solver/fem_native.py:3311-3318).
All work runs in the HIP library; this module only allocates and wires device buffers.
"""
from __future__ import annotations

import ctypes as C
import threading
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Optional, Union

import torch

from . import _lib
from .mesh import Mesh

# Plans keyed on the content of the cells (and the sizes and vec), like compress_plan's memo
# (jno/utils/solver/fem_utils.py:4027-4060): a mesh rebuilt with the same connectivity (C4's many
# boxes of one size, a re-created mesh) reuses the plan instead of building it again; a remeshed
# domain changes the content and misses.  Least recently used plans are dropped past the budget;
# the lock makes lookups safe from several threads.
PLAN_CACHE_BYTES = 16 << 30
PLAN_CACHE_MAX = 256
_PLAN_CACHE: "OrderedDict" = OrderedDict()
_PLAN_LOCK = threading.Lock()


def _cells_key(mesh: Mesh, vec: int):
    ws = torch.empty(256, dtype=torch.uint8, device=mesh.device)
    h = (C.c_uint64 * 2)()
    _lib.call("jno_hash_i32", _lib.ptr(mesh.cells), mesh.cells.numel(), h, _lib.ptr(ws), 256,
              _lib.stream_ptr(mesh.device))
    dev = mesh.device.index if mesh.device.index is not None else torch.cuda.current_device()
    return (dev, int(h[0]), int(h[1]), mesh.n_cells, mesh.nloc, mesh.n_nodes, vec)


def _same_cells(a: torch.Tensor, b: torch.Tensor) -> bool:
    if a is b:
        return True
    if a.shape != b.shape:
        return False
    ws = torch.empty(256, dtype=torch.uint8, device=a.device)
    eq = C.c_int(0)
    _lib.call("jno_equal_i32", _lib.ptr(a), _lib.ptr(b), a.numel(), C.byref(eq), _lib.ptr(ws), 256,
              _lib.stream_ptr(a.device))
    return bool(eq.value)


def _cache_bytes() -> int:
    plans = {id(p): p.nbytes for p, _ in _PLAN_CACHE.values()}
    return sum(plans.values()) + sum(c.numel() * 4 for _, c in _PLAN_CACHE.values() if c is not None)


def plan_for(mesh: Mesh, vec: int = 1) -> "Plan":
    """The plan of ``mesh`` for ``vec`` components, from the cache or built.  A structured Kuhn box
    (checked once per mesh, :meth:`BoxAssembly.of`) gets the closed-form plan keyed on its grid;
    any other mesh the sort-based one keyed on the content of its cells."""
    box = BoxAssembly.of(mesh) if vec == 1 and mesh.nloc == 4 and mesh.grid is not None else None
    if box is not None:
        key = ("kuhn_box", box.grid, str(mesh.device))
        check = None
    else:
        key = _cells_key(mesh, vec)
        check = mesh.cells
    with _PLAN_LOCK:
        hit = _PLAN_CACHE.get(key)
        if hit is not None and (check is None or _same_cells(hit[1], check)):
            _PLAN_CACHE.move_to_end(key)
            return hit[0]
    plan = Plan.build_kuhn_box(mesh, box.grid) if box is not None else Plan.build(mesh, vec)
    # a box plan is also filed under the content of its cells, so that a mesh with the same
    # connectivity and no lattice (moved points, or built from the same tensors) shares it
    entries = [(key, (plan, check))] + ([(_cells_key(mesh, vec), (plan, mesh.cells))] if box is not None else [])
    with _PLAN_LOCK:
        for k_, e in entries:
            _PLAN_CACHE[k_] = e
            _PLAN_CACHE.move_to_end(k_)
        while len(_PLAN_CACHE) > 2 and (len(_PLAN_CACHE) > PLAN_CACHE_MAX or _cache_bytes() > PLAN_CACHE_BYTES):
            _PLAN_CACHE.popitem(last=False)
    return plan


def clear_plan_cache():
    with _PLAN_LOCK:
        _PLAN_CACHE.clear()


@dataclass
class Plan:
    """CSR index structure (the bit-exact contract) + the node->(element, local) incidence.

    ``rowptr``/``col`` equal ``compress_plan``'s (row, col)-sorted unique pairs
    (jno/utils/solver/fem_utils.py:4009-4060) for dof = node*vec + comp."""

    rowptr: torch.Tensor
    col: torch.Tensor
    inc_ptr: torch.Tensor
    inc: torch.Tensor
    nnz: int
    max_row: int
    ndof: int
    vec: int

    @staticmethod
    def build(mesh: Mesh, vec: int = 1) -> "Plan":
        dev = mesh.device
        N, T, nloc = mesh.n_nodes, mesh.n_cells, mesh.nloc
        ndof = N * vec
        inc_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
        inc = torch.empty(max(T * nloc, 1), dtype=torch.int32, device=dev)
        rowptr = torch.empty(ndof + 1, dtype=torch.int32, device=dev)
        ws_bytes = _lib.load().jno_pattern_workspace_size(T, nloc, N)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        nnz = C.c_int64(0)
        max_row = C.c_int32(0)
        stream = _lib.stream_ptr(dev)
        _lib.call(
            "jno_pattern_count", _lib.ptr(mesh.cells), T, nloc, vec, N, _lib.ptr(inc_ptr), _lib.ptr(inc),
            _lib.ptr(rowptr), C.byref(nnz), C.byref(max_row), _lib.ptr(ws), ws_bytes, stream,
        )
        col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
        _lib.call(
            "jno_pattern_fill", _lib.ptr(mesh.cells), T, nloc, vec, N, _lib.ptr(inc_ptr), _lib.ptr(inc),
            _lib.ptr(rowptr), _lib.ptr(col), stream,
        )
        return Plan(rowptr, col[: nnz.value], inc_ptr, inc, int(nnz.value), int(max_row.value), ndof, vec)

    @staticmethod
    def build_kuhn_box(mesh: Mesh, grid) -> "Plan":
        """The same plan for the structured Kuhn box ``grid`` in closed form (one pass, no sort)."""
        dev = mesh.device
        nx, ny, nz = grid
        N, T = mesh.n_nodes, mesh.n_cells
        nnz = C.c_int64(0)
        max_row = C.c_int32(0)
        _lib.call("jno_pattern_kuhn_box_nnz", nx, ny, nz, C.byref(nnz), C.byref(max_row))
        rowptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
        inc_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
        col = torch.empty(nnz.value, dtype=torch.int32, device=dev)
        inc = torch.empty(4 * T, dtype=torch.int32, device=dev)
        _lib.call("jno_pattern_kuhn_box", nx, ny, nz, _lib.ptr(rowptr), _lib.ptr(col), _lib.ptr(inc_ptr),
                  _lib.ptr(inc), _lib.stream_ptr(dev))
        return Plan(rowptr, col, inc_ptr, inc, int(nnz.value), int(max_row.value), N, 1)

    @property
    def nbytes(self) -> int:
        return int(sum(t.numel() * t.element_size() for t in (self.rowptr, self.col, self.inc_ptr, self.inc)))


@dataclass
class CSR:
    """An assembled sparse operator in CSR form (FP64 values) on the device.

    Plays the role of the reference's ``LinearOperator`` over a BCOO
    (jno/utils/solver/solver_api.py:82-184): ``mv``/``@``, ``diag``, ``T`` (symmetric operators
    are their own transpose here), ``dense`` for small checks, and ``bcoo`` = (data, idx)."""

    rowptr: torch.Tensor
    col: torch.Tensor
    vals: torch.Tensor
    shape: tuple
    dinv: Optional[torch.Tensor] = None  # Jacobi inverse diagonal, when assembly produced it
    _cache: dict = field(default_factory=dict, repr=False)
    # operator family on a shared pattern ("" for the assembled operator; the transient step
    # operator uses its own): layouts learnt on one family are not offered to another
    family: str = ""

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    @property
    def device(self):
        return self.rowptr.device

    def _spmv(self, x: torch.Tensor) -> torch.Tensor:
        x = x.detach().to(torch.float64).contiguous()
        y = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        _lib.call("jno_spmv_f64", self.shape[0], _lib.ptr(self.rowptr), _lib.ptr(self.col), _lib.ptr(self.vals),
                  _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr(self.device))
        return y

    def mv(self, x: torch.Tensor) -> torch.Tensor:
        """y = A x; differentiable in x (the operators assembled here are symmetric, so the
        cotangent is A g)."""
        if torch.is_grad_enabled() and x.requires_grad:
            return _MatVec.apply(x, self)
        return self._spmv(x)

    __matmul__ = mv
    __call__ = mv

    @property
    def T(self) -> "CSR":
        return self

    def diag_inv(self) -> torch.Tensor:
        """Jacobi inverse diagonal with the |d| > 1e-30 guard (jno/precond.py:105-109)."""
        if self.dinv is None:
            d = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
            _lib.call("jno_jacobi", self.shape[0], _lib.ptr(self.rowptr), _lib.ptr(self.col), _lib.ptr(self.vals),
                      _lib.ptr(d), _lib.stream_ptr(self.device))
            self.dinv = d
        return self.dinv

    def diag(self) -> torch.Tensor:
        """The diagonal (duplicates summed, ``matrix_diagonal``, jno/utils/solver/linear.py:88-94)."""
        rows = self.row_indices()
        on = rows == self.col.long()
        d = torch.zeros(self.shape[0], dtype=torch.float64, device=self.device)
        d.index_add_(0, rows[on], self.vals[on])
        return d

    def row_indices(self) -> torch.Tensor:
        counts = (self.rowptr[1:] - self.rowptr[:-1]).long()
        return torch.repeat_interleave(torch.arange(self.shape[0], device=self.device), counts)

    def dense(self) -> torch.Tensor:
        Dm = torch.zeros(self.shape, dtype=torch.float64, device=self.device)
        Dm.index_put_((self.row_indices(), self.col.long()), self.vals, accumulate=True)
        return Dm

    @property
    def bcoo(self):
        """(data, idx) with idx (nnz, 2) int32 sorted by (row, col), the BCOO view."""
        return self.vals, torch.stack([self.row_indices().to(torch.int32), self.col], dim=1)

    def to_bcoo(self) -> BCOO:
        """The reference's BCOO of this operator (exact zeros dropped, sorted row-major)."""
        Az = self.drop_zeros()
        data, idx = Az.bcoo
        return BCOO(data, idx.contiguous(), self.shape)

    def drop_zeros(self) -> "CSR":
        """``compress_eager`` (jno/utils/solver/fem_utils.py:4063-4091): entries exactly 0.0 removed."""
        n = self.shape[0]
        rowptr_o = torch.empty(n + 1, dtype=torch.int32, device=self.device)
        ws_bytes = (n + 1) * 4 + 512 + ((n + 1023) // 1024 + 1) * 8 + 256
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        nnz = C.c_int64(0)
        stream = _lib.stream_ptr(self.device)
        _lib.call("jno_drop_zeros_count", n, _lib.ptr(self.rowptr), _lib.ptr(self.vals), _lib.ptr(rowptr_o),
                  C.byref(nnz), _lib.ptr(ws), ws_bytes, stream)
        col_o = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=self.device)
        vals_o = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=self.device)
        _lib.call("jno_drop_zeros_fill", n, _lib.ptr(self.rowptr), _lib.ptr(self.col), _lib.ptr(self.vals),
                  _lib.ptr(rowptr_o), _lib.ptr(col_o), _lib.ptr(vals_o), stream)
        return CSR(rowptr_o, col_o[: nnz.value], vals_o[: nnz.value], self.shape, self.dinv)


@dataclass
class BCOO:
    """The assembled operator as the reference stores it: a ``jax.experimental.sparse.BCOO`` after
    ``compress_eager`` (jno/utils/solver/fem_utils.py:4063-4091; docs/fem.md:974-983): one entry per
    (row, col), sorted row-major, exact zeros dropped (P_eager).  ``data`` (nse,) FP64 and
    ``indices`` (nse, 2) int32 on the device, byte-for-byte the reference's index array."""

    data: torch.Tensor
    indices: torch.Tensor
    shape: tuple

    @property
    def nse(self) -> int:
        return int(self.data.shape[0])

    def todense(self) -> torch.Tensor:
        Dm = torch.zeros(self.shape, dtype=torch.float64, device=self.data.device)
        Dm.index_put_((self.indices[:, 0].long(), self.indices[:, 1].long()), self.data, accumulate=True)
        return Dm


class _MatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, A: CSR):
        ctx.A = A
        return A._spmv(x)

    @staticmethod
    def backward(ctx, gy):
        return ctx.A._spmv(gy), None


# Kuhn box stencil (the kernel's compile-time table): the 15 offsets (di, dj, dk) of a node's
# couplings in ascending node order, and the (corner, type, local vertex) of its incident cells
_KUHN = ((0, 4, 6, 7), (0, 4, 5, 7), (0, 2, 6, 7), (0, 2, 3, 7), (0, 1, 5, 7), (0, 1, 3, 7))


class BoxAssembly:
    """What the structured-box assembly needs of a mesh (built once per mesh): the six cell types'
    geometry table and, for the stiffness, the DIA offsets that can hold a nonzero coupling."""

    def __init__(self, mesh: Mesh, geo):
        nx, ny, nz = mesh.grid
        self.grid = (nx, ny, nz)
        self.geo = geo
        S2, S1 = nz + 1, (ny + 1) * (nz + 1)
        offs = {0}
        for t in range(6):
            for a in range(4):
                c = _KUHN[t][a]
                for bb in range(4):
                    cb = _KUHN[t][bb]
                    if geo[6 + (t * 4 + a) * 4 + bb] != 0.0:
                        d = (((cb >> 2) & 1) - ((c >> 2) & 1), ((cb >> 1) & 1) - ((c >> 1) & 1), (cb & 1) - (c & 1))
                        offs.add(abs(d[0] * S1 + d[1] * S2 + d[2]))
        self.stiffness_offsets = sorted(offs)

    def covers(self, offsets) -> bool:
        return set(self.stiffness_offsets) <= set(offsets)

    @staticmethod
    def of(mesh: Mesh) -> Optional["BoxAssembly"]:
        """The box assembly of ``mesh`` when it is (bitwise) the structured Kuhn box of its grid and
        box with at least 2 cells per axis, else None.  Checked once (one stream synchronization)
        and kept on the mesh with the points tensor's identity and version: moving the points with
        torch (``mesh.points.copy_(..)``, a new tensor) checks again, and a mesh found off the lattice
        stays on the generic assembly.  Points written behind torch's back (raw pointers) need
        ``mesh.invalidate()``."""
        key = ("box_assembly",)
        pts = mesh.points
        stamp = (id(pts), pts.data_ptr(), pts._version)
        hit = mesh._cache.get(key)
        if hit is not None and (hit[0] is None or hit[1] == stamp):
            return hit[0]
        res = None
        if mesh.grid is not None and mesh.box is not None and mesh.nloc == 4 and min(mesh.grid) >= 2:
            nx, ny, nz = mesh.grid
            ws = torch.empty(1024, dtype=torch.uint8, device=mesh.device)
            ok = C.c_int(0)
            lo = (C.c_double * 3)(*mesh.box[0])
            hi = (C.c_double * 3)(*mesh.box[1])
            stream = _lib.stream_ptr(mesh.device)
            _lib.call("jno_mesh_is_kuhn_box", nx, ny, nz, lo, hi, _lib.ptr(mesh.points), _lib.ptr(mesh.cells),
                      C.byref(ok), _lib.ptr(ws), 1024, stream)
            if ok.value:
                geo = (C.c_double * (6 * 17))()
                _lib.call("jno_box_geometry", nx, ny, nz, _lib.ptr(mesh.points), geo, _lib.ptr(ws), 1024, stream)
                res = BoxAssembly(mesh, geo)
        mesh._cache[key] = (res, stamp)
        return res


def _as_k(k, T: int, device) -> torch.Tensor:
    k = torch.as_tensor(k, dtype=torch.float64, device=device)
    if k.ndim == 0:
        k = k.expand(T)
    if k.shape != (T,):
        raise ValueError(f"the per-element coefficient must have shape ({T},) (one value per cell), got {tuple(k.shape)}")
    return k.contiguous()


class FemLinearSystem:
    """-div(k grad u) = f on a P1 tet mesh with u = g on ``dirichlet``; ``k`` is a per-element (P0)
    runtime parameter named ``param``.

    ``evaluate({"k": k})`` re-assembles ``A(k), b(k)`` like the reference's parametric operator
    (jno/trace/__init__.py:3612-3615): CSR on the structural pattern P_struct, with the symmetric
    Dirichlet elimination of jno/utils/solver/fem_1d.py:667-686 fused into assembly.
    """

    def __init__(self, mesh: Mesh, f: Union[float, torch.Tensor] = 1.0, dirichlet: Optional[str] = "boundary",
                 g: Union[float, torch.Tensor] = 0.0, param: str = "k", structured: bool = True):
        if mesh.nloc != 4:
            raise ValueError("FemLinearSystem (P1 Poisson) needs a 4-node tet mesh")
        self.mesh = mesh
        self.param = param
        # the structured-box kernel on a Kuhn box mesh (False: always the incidence-gather kernel)
        self.structured = bool(structured)
        self.plan = mesh.plan(1)
        N = mesh.n_nodes
        dev = mesh.device
        self.f_const = 0.0
        self.f_nodal = None
        if isinstance(f, torch.Tensor) and f.ndim > 0:
            if f.shape != (N,):
                raise ValueError(f"nodal load must have shape ({N},)")
            self.f_nodal = f.to(device=dev, dtype=torch.float64).contiguous()
        else:
            self.f_const = float(f)
        self.dmask = mesh.boundary_mask(dirichlet) if dirichlet else None
        self.g = None
        if isinstance(g, torch.Tensor) and g.ndim > 0:
            self.g = g.to(device=dev, dtype=torch.float64).contiguous()
        elif float(g) != 0.0:
            self.g = torch.full((N,), float(g), dtype=torch.float64, device=dev)

    @property
    def n(self) -> int:
        return self.mesh.n_nodes

    def evaluate(self, args: dict, out: Optional[CSR] = None):
        """(A, b) at the parameter values ``args[param]`` (shape (T,))."""
        if self.param not in args:
            raise KeyError(f"evaluate() needs the runtime parameter {self.param!r}")
        m, plan = self.mesh, self.plan
        k = _as_k(args[self.param], m.n_cells, m.device)
        N = m.n_nodes
        if out is None:
            vals = torch.empty(plan.nnz, dtype=torch.float64, device=m.device)
            dinv = torch.empty(N, dtype=torch.float64, device=m.device)
            A = CSR(plan.rowptr, plan.col, vals, (N, N), dinv)
        else:
            A = out
            A._cache.clear()
        b = torch.empty(N, dtype=torch.float64, device=m.device)
        from .operator import SolverOperator, chosen_dia_offsets

        dia = chosen_dia_offsets(A)
        box = BoxAssembly.of(m) if self.structured else None
        if box is not None:
            # the structured box: no incidence or geometry gathers, and the DIA layout (when the
            # learnt offsets cover every nonzero coupling the geometry table allows) in the same pass
            # with no flag to read back
            use_dia = dia is not None and box.covers(dia[1])
            nd, offs = dia if use_dia else (0, [0])
            ld = (N + 63) // 64 * 64
            d64 = torch.empty(nd * ld, dtype=torch.float64, device=m.device) if use_dia else None
            d32 = torch.empty(nd * ld, dtype=torch.float32, device=m.device) if use_dia else None
            nx, ny, nz = box.grid
            _lib.call(
                "jno_assemble_p1_box", 0, nx, ny, nz, box.geo, _lib.ptr(plan.rowptr), _lib.ptr(k), self.f_const,
                _lib.ptr(self.f_nodal), _lib.ptr(self.dmask), _lib.ptr(self.g), _lib.ptr(A.vals), _lib.ptr(b),
                _lib.ptr(A.dinv), nd, (C.c_int32 * _lib.JNO_DIA_MAX)(*offs), _lib.ptr(d64), _lib.ptr(d32),
                _lib.stream_ptr(m.device),
            )
            if use_dia:
                op = SolverOperator.from_dia(A, nd, offs, d64, d32)
                A._cache[("solver_op", "auto", True)] = op
                A._cache[("solver_op", "auto", False)] = op
            return A, b
        if dia is None:
            _lib.call(
                "jno_assemble_p1", 0, _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, N, _lib.ptr(plan.inc_ptr),
                _lib.ptr(plan.inc), _lib.ptr(plan.rowptr), _lib.ptr(plan.col), plan.max_row, _lib.ptr(k), self.f_const,
                _lib.ptr(self.f_nodal), _lib.ptr(self.dmask), _lib.ptr(self.g), _lib.ptr(A.vals), _lib.ptr(b),
                _lib.ptr(A.dinv), _lib.stream_ptr(m.device),
            )
            return A, b
        # the solver's DIA layout (offsets found on an earlier operator of this plan) in the same pass
        nd, offs = dia
        ld = (N + 63) // 64 * 64
        d64 = torch.empty(nd * ld, dtype=torch.float64, device=m.device)
        d32 = torch.empty(nd * ld, dtype=torch.float32, device=m.device)
        ws = torch.empty(256, dtype=torch.uint8, device=m.device)
        ok = C.c_int(0)
        _lib.call(
            "jno_assemble_p1_dia", 0, _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, N, _lib.ptr(plan.inc_ptr),
            _lib.ptr(plan.inc), _lib.ptr(plan.rowptr), _lib.ptr(plan.col), plan.max_row, _lib.ptr(k), self.f_const,
            _lib.ptr(self.f_nodal), _lib.ptr(self.dmask), _lib.ptr(self.g), _lib.ptr(A.vals), _lib.ptr(b),
            _lib.ptr(A.dinv), nd, (C.c_int32 * _lib.JNO_DIA_MAX)(*offs), _lib.ptr(d64), _lib.ptr(d32),
            C.byref(ok), _lib.ptr(ws), 256, _lib.stream_ptr(m.device),
        )
        if ok.value:
            op = SolverOperator.from_dia(A, nd, offs, d64, d32)
            A._cache[("solver_op", "auto", True)] = op
            A._cache[("solver_op", "auto", False)] = op
        return A, b

    def assembled(self, args: dict):
        """(A, b) as the reference's non-parametric ``fem.A`` / ``fem.b`` hold them: A as a
        :class:`BCOO` on P_eager (docs/fem.md:964-983), b the eliminated load."""
        A, b = self.evaluate(args)
        return A.to_bcoo(), b

    def mass(self) -> CSR:
        """Consistent P1 mass matrix (the u*v form) on the same pattern, no Dirichlet elimination."""
        m, plan = self.mesh, self.plan
        N = m.n_nodes
        vals = torch.empty(plan.nnz, dtype=torch.float64, device=m.device)
        box = BoxAssembly.of(m) if self.structured else None
        if box is not None:
            nx, ny, nz = box.grid
            _lib.call("jno_assemble_p1_box", 1, nx, ny, nz, box.geo, _lib.ptr(plan.rowptr), None, 0.0, None, None,
                      None, _lib.ptr(vals), None, None, 0, (C.c_int32 * _lib.JNO_DIA_MAX)(0), None, None,
                      _lib.stream_ptr(m.device))
            return CSR(plan.rowptr, plan.col, vals, (N, N))
        _lib.call(
            "jno_assemble_p1", 1, _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, N, _lib.ptr(plan.inc_ptr),
            _lib.ptr(plan.inc), _lib.ptr(plan.rowptr), _lib.ptr(plan.col), plan.max_row, None, 0.0, None, None, None,
            _lib.ptr(vals), None, None, _lib.stream_ptr(m.device),
        )
        return CSR(plan.rowptr, plan.col, vals, (N, N))

    def grad_k(self, u: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
        """dL/dk_e = -(P lam)_e^T (K_e/k_e) u_e for the adjoint state ``lam`` (P zeroes Dirichlet dofs)."""
        m = self.mesh
        grad = torch.empty(m.n_cells, dtype=torch.float64, device=m.device)
        u = u.to(torch.float64).contiguous()
        lam = lam.to(torch.float64).contiguous()
        _lib.call("jno_adjoint_grad_k_p1", _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, _lib.ptr(u), _lib.ptr(lam),
                  _lib.ptr(self.dmask), _lib.ptr(grad), _lib.stream_ptr(m.device))
        return grad


def fem_poisson(mesh: Mesh, f=1.0, dirichlet="boundary", g=0.0, param="k", structured=True) -> FemLinearSystem:
    """Spelling of ``jno.fem([k*(grad u . grad v) - f*v, u(boundary) - g])`` with ``k`` a P0 parameter."""
    return FemLinearSystem(mesh, f=f, dirichlet=dirichlet, g=g, param=param, structured=structured)
