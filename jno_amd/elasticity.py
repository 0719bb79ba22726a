"""P2 Lagrange 3-D linear elasticity (10-node tets, 3 dofs per node) with a per-element Young's
modulus ``E`` as the runtime parameter, clamped on one box face and loaded by a constant traction
on another.

Numeric analogue of ``jno.fem([E*(lam tr eps(u) tr eps(v) + 2 mu eps(u):eps(v)), u(clamp) - 0,
-t . v(load)])`` with ``E`` a P0 ``jno.np.parameter`` (weak form as in
tests/test_topology_optimisation.py:59-66, surface term jno/utils/solver/fem_native.py:1882-1960).
All work runs in the HIP library: P2 promotion, coloring, MFMA element matrices with a colored
scatter, traction load, Dirichlet elimination and the compliance gradient.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import _lib
from .fem import CSR, Plan
from .mesh import FACES, Mesh

MAX_COLORS = 128


def promote_p2(mesh: Mesh) -> Mesh:
    """The P2 node mesh of a P1 tet mesh (``_promote_to_degree``, fem_utils.py:1979-2044)."""
    if mesh.order != 1 or mesh.nloc != 4:
        raise ValueError("promote_p2 needs a P1 (4-node) tet mesh")
    key = ("p2",)
    if key in mesh._cache:
        return mesh._cache[key]
    plan = mesh.plan(1)
    T, N = mesh.n_cells, mesh.n_nodes
    dev = mesh.device
    cells10 = torch.empty((T, 10), dtype=torch.int32, device=dev)
    pts2 = torch.empty((N + 6 * T, 3), dtype=torch.float64, device=dev)
    wsb = _lib.load().jno_promote_p2_workspace_size(T)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    n2 = C.c_int64(0)
    _lib.call("jno_promote_p2", _lib.ptr(mesh.points), _lib.ptr(mesh.cells), T, N, _lib.ptr(plan.inc_ptr),
              _lib.ptr(plan.inc), _lib.ptr(cells10), _lib.ptr(pts2), C.byref(n2), _lib.ptr(ws), wsb,
              _lib.stream_ptr(dev))
    m2 = Mesh(pts2[: n2.value].contiguous(), cells10, 2, mesh.box, mesh.grid)
    mesh._cache[key] = m2
    return m2


class Coloring:
    """Cells grouped so that no two cells of a group share a node."""

    def __init__(self, mesh: Mesh, plan: Plan):
        T = mesh.n_cells
        self.order = torch.empty(max(T, 1), dtype=torch.int32, device=mesh.device)
        self.color_ptr = (C.c_int32 * (MAX_COLORS + 1))()
        nc = C.c_int32(0)
        wsb = _lib.load().jno_color_workspace_size(T)
        ws = torch.empty(wsb, dtype=torch.uint8, device=mesh.device)
        _lib.call("jno_color_cells", _lib.ptr(mesh.cells), T, mesh.nloc, _lib.ptr(plan.inc_ptr), _lib.ptr(plan.inc),
                  _lib.ptr(self.order), self.color_ptr, MAX_COLORS, C.byref(nc), _lib.ptr(ws), wsb,
                  _lib.stream_ptr(mesh.device))
        self.ncolors = int(nc.value)

    def groups(self):
        return [(self.color_ptr[c], self.color_ptr[c + 1]) for c in range(self.ncolors)]


class ElasticitySystem:
    """``evaluate({"E": E})`` -> (A, b) on the P2 vector pattern, u = 0 on ``clamp`` (all three
    components), constant ``traction`` on ``load``."""

    def __init__(self, mesh: Mesh, nu: float = 0.3, traction=(0.0, 0.0, -1.0), clamp: str = "left",
                 load: str = "right", param: str = "E"):
        if not (-1.0 < nu < 0.5):
            raise ValueError("the Poisson ratio must lie in (-1, 0.5)")
        self.p1 = mesh
        self.mesh = promote_p2(mesh) if mesh.order == 1 else mesh
        m = self.mesh
        self.nu = float(nu)
        self.traction = tuple(float(t) for t in traction)
        self.param = param
        self.plan = m.plan(3)
        self.coloring = Coloring(m, self.plan)
        # each (cell, node pair)'s place in its CSR row, found once, so the assembly scatters with no
        # searches (None when the table does not apply: the assembly searches)
        self.pair_rank = torch.empty(max(100 * m.n_cells, 1), dtype=torch.uint8, device=m.device)
        ok = C.c_int32(0)
        ws = torch.empty(256, dtype=torch.uint8, device=m.device)
        _lib.call("jno_p2_pair_rank", _lib.ptr(m.cells), m.n_cells, _lib.ptr(self.plan.rowptr), _lib.ptr(self.plan.col),
                  _lib.ptr(self.pair_rank), C.byref(ok), _lib.ptr(ws), 256, _lib.stream_ptr(m.device))
        if not ok.value:
            self.pair_rank = None
        # assemble from the geometry-class table when the cells repeat their shape (row-pull, no
        # element matrices per cell); False: MFMA element matrices and the colored scatter
        self.class_assembly = True
        self.ndof = 3 * m.n_nodes
        node_mask = m.boundary_mask(clamp) if clamp else None
        self.dmask = node_mask.repeat_interleave(3).contiguous() if node_mask is not None else None
        axis, end = FACES[load]
        self.load_axis, self.load_value = axis, float(m.box[end][axis])
        # the load does not depend on E: build it once
        self._b0 = torch.zeros(self.ndof, dtype=torch.float64, device=m.device)
        t = (C.c_double * 3)(*self.traction)
        _lib.call("jno_traction_p2", _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, _lib.ptr(self.plan.inc_ptr),
                  _lib.ptr(self.plan.inc), axis, self.load_value, t, _lib.ptr(self._b0), _lib.stream_ptr(m.device))

    def traction_load(self) -> torch.Tensor:
        """The assembled traction load int_F t . phi_i dS (3 dofs per node, before the Dirichlet
        elimination); it does not depend on E."""
        return self._b0

    def evaluate(self, args: dict, out: Optional[CSR] = None):
        m, plan = self.mesh, self.plan
        E = torch.as_tensor(args[self.param], dtype=torch.float64, device=m.device)
        if E.ndim == 0:
            E = E.expand(m.n_cells)
        if E.shape != (m.n_cells,):
            raise ValueError(f"{self.param} must have one value per cell ({m.n_cells},)")
        E = E.contiguous()
        n = self.ndof
        if out is None:
            vals = torch.empty(plan.nnz, dtype=torch.float64, device=m.device)
            A = CSR(plan.rowptr, plan.col, vals, (n, n))
        else:
            A = out
            A._cache.clear()
            A.dinv = None
        stream = _lib.stream_ptr(m.device)
        geo = self._class_geometry() if self.class_assembly and self.pair_rank is not None else None
        b = self._b0.clone()
        if geo is not None and plan.max_row <= 384:
            # the class-table row pull with the elimination fused (zero Dirichlet values: b[d] = 0)
            _lib.call("jno_assemble_p2_classes", m.n_nodes, _lib.ptr(plan.inc_ptr), _lib.ptr(plan.inc),
                      _lib.ptr(m.cells), _lib.ptr(geo.cell_class), _lib.ptr(geo.table64), _lib.ptr(E),
                      _lib.ptr(self.pair_rank), _lib.ptr(self.dmask), _lib.ptr(plan.rowptr), plan.max_row,
                      _lib.ptr(A.vals), stream)
            if self.dmask is not None:
                _lib.call("jno_ebe_dirichlet_rhs", n, _lib.ptr(self.dmask), _lib.ptr(b), stream)
            return A, b
        _lib.call("jno_assemble_p2_elasticity", _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells,
                  _lib.ptr(plan.rowptr), _lib.ptr(plan.col), plan.nnz, _lib.ptr(E), self.nu,
                  _lib.ptr(self.coloring.order), self.coloring.color_ptr, self.coloring.ncolors,
                  _lib.ptr(self.pair_rank), _lib.ptr(A.vals), stream)
        if self.dmask is not None:
            _lib.call("jno_dirichlet_eliminate", n, _lib.ptr(plan.rowptr), _lib.ptr(plan.col), _lib.ptr(A.vals),
                      _lib.ptr(b), _lib.ptr(self.dmask), None, stream)
        return A, b

    def _class_geometry(self):
        """The geometry classes of the cells (shared with the element-by-element operator), or None
        when the cells do not repeat their shape."""
        from .ebe import EbeGeometry

        geo = getattr(self, "_ebe_geometry", None)
        if geo is None and not getattr(self, "_no_classes", False):
            try:
                geo = EbeGeometry(self)
                self._ebe_geometry = geo
            except ValueError:
                self._no_classes = True
        return geo

    def element_operator(self, args: dict, structured: bool = True):
        """(A, b) with A the element-by-element operator (:class:`jno_amd.ebe.ElementOperator`, the
        reference's ElementOperator, jno/utils/solver/element_operator.py:1-178) instead of the
        assembled CSR: the same eliminated operator and load, no global matrix.  On a structured
        Kuhn box its FP32 products run on the half lattice (``structured=False``: the general
        element-by-element kernels)."""
        from .ebe import ElementOperator

        m = self.mesh
        E = torch.as_tensor(args[self.param], dtype=torch.float64, device=m.device)
        if E.ndim == 0:
            E = E.expand(m.n_cells)
        if E.shape != (m.n_cells,):
            raise ValueError(f"{self.param} must have one value per cell ({m.n_cells},)")
        A = ElementOperator(self, E.contiguous(), structured=structured)
        b = self._b0.clone()
        if self.dmask is not None:
            _lib.call("jno_ebe_dirichlet_rhs", self.ndof, _lib.ptr(self.dmask), _lib.ptr(b), _lib.stream_ptr(m.device))
        return A, b

    def compliance_grad(self, u: torch.Tensor) -> torch.Tensor:
        """dC/dE_e = -u_e^T (K_e/E_e) u_e for C = b.u."""
        m = self.mesh
        g = torch.empty(m.n_cells, dtype=torch.float64, device=m.device)
        u = u.to(torch.float64).contiguous()
        _lib.call("jno_compliance_grad_p2", _lib.ptr(m.points), _lib.ptr(m.cells), m.n_cells, self.nu, _lib.ptr(u),
                  _lib.ptr(g), _lib.stream_ptr(m.device))
        return g


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, system: ElasticitySystem, solver, operator: str):
        from .precond import jacobi

        if operator == "ebe":  # the element-by-element operator: no assembled matrix
            A, b = system.element_operator({system.param: E.detach()})
        else:
            A, b = system.evaluate({system.param: E.detach()})
        u = solver(A, b, M=jacobi())
        ctx.save_for_backward(u)
        ctx.system = system
        return torch.dot(b, u), u

    @staticmethod
    def backward(ctx, gC, gu):
        (u,) = ctx.saved_tensors
        if gu is not None and bool((gu != 0).any()):
            raise NotImplementedError("only the compliance output is differentiable here")
        return gC * ctx.system.compliance_grad(u), None, None, None


def compliance(system: ElasticitySystem, E: torch.Tensor, solver=None, operator: str = "ebe"):
    """(C, u) with C = b.u differentiable with respect to the per-element modulus ``E``.  The forward
    solve runs on the element-by-element operator (``operator="ebe"``, the default: 2.7x faster
    than the assembled BSR3 layout at 32^3) or on the assembled matrix (``"assembled"``); the
    gradient is the closed-form contraction dC/dE_e = -u_e^T (K_e/E_e) u_e either way."""
    from .solve import cg

    if operator not in ("ebe", "assembled"):
        raise ValueError("operator must be 'ebe' or 'assembled'")
    return _Compliance.apply(E, system, solver or cg(tol=1e-10), operator)
