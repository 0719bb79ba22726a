"""Tetrahedral meshes on the device: the structured Kuhn box of ``jno.Shape.box(..).structured(n)``.

Mirrors the reference spelling ``jno.Shape.box(x0, y0, z0, x1, y1, z1).structured(n).domain()``
(jno/geometry/shape.py:415-458 -> jno/domain/domain_class.py:2259-2330 ->
Geometries.equi_distant_box, jno/domain/geometries.py:204-301).  The lattice is generated by the
``jno_mesh_kuhn_box`` HIP kernel straight into HBM.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Tuple

import torch

from . import _lib

FACES = {  # named box faces: (axis, which end)
    "left": (0, 0),
    "right": (0, 1),
    "bottom": (1, 0),
    "top": (1, 1),
    "front": (2, 0),
    "back": (2, 1),
}


@dataclass
class Mesh:
    """A conforming tet mesh: ``points`` (N,3) f64 and ``cells`` (T, 4|10) int32 on the device.

    ``cells`` columns follow basix DOF order (vertices, then P2 edge nodes on edges
    (2,3),(1,3),(1,2),(0,3),(0,2),(0,1), jno/utils/solver/fem_topology.py:44).  ``box`` is the
    (lo, hi) corners when the mesh is a structured box (named faces resolve against it)."""

    points: torch.Tensor
    cells: torch.Tensor
    order: int = 1
    box: Optional[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = None
    grid: Optional[Tuple[int, int, int]] = None
    _cache: dict = field(default_factory=dict, repr=False)

    @property
    def n_nodes(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def nloc(self) -> int:
        return int(self.cells.shape[1])

    @property
    def device(self):
        return self.points.device

    def centroids(self) -> torch.Tensor:
        """Element centroids (mean of the 4 vertices), (T,3) f64."""
        v = self.points[self.cells[:, :4].long()]
        return v.mean(dim=1)

    def plan(self, vec: int = 1):
        """The CSR pattern + incidence plan for ``vec`` components: from the content-keyed plan cache
        (``jno_amd.fem.plan_for``; a mesh with the same cells shares it), kept on the mesh."""
        from .fem import plan_for

        key = ("plan", vec)
        if key not in self._cache:
            self._cache[key] = plan_for(self, vec)
        return self._cache[key]

    def invalidate(self) -> None:
        """Forget everything derived from the points and cells (plans, masks, the P2 promotion, the
        structured-box check): call it after changing them in place."""
        self._cache.clear()

    def boundary_mask(self, region: str = "boundary") -> torch.Tensor:
        """u8[N] mask of the nodes of ``region``: "boundary" = every node on a facet used by
        exactly one cell (``_boundary_facets``, jno/_fem.py:2695-2773); a named box face keeps the
        boundary nodes whose coordinate lies on that face (``_boundary_node_ids``,
        jno/utils/solver/fem_native.py:2563-2636)."""
        key = ("bmask", region)
        if key in self._cache:
            return self._cache[key]
        stream = _lib.stream_ptr(self.device)
        allb = self._cache.get(("bmask", "boundary"))
        from .fem import BoxAssembly

        box = BoxAssembly.of(self) if allb is None and self.grid is not None and self.nloc == 4 else None
        if box is not None:  # the structured Kuhn box: its surface nodes in closed form
            allb = torch.empty(self.n_nodes, dtype=torch.uint8, device=self.device)
            _lib.call("jno_mask_kuhn_box", *box.grid, _lib.ptr(allb), stream)
            self._cache[("bmask", "boundary")] = allb
        if allb is None:
            plan = self.plan(1)
            allb = torch.empty(self.n_nodes, dtype=torch.uint8, device=self.device)
            _lib.call(
                "jno_boundary_mask",
                _lib.ptr(self.cells),
                self.n_cells,
                self.nloc,
                self.n_nodes,
                _lib.ptr(plan.inc_ptr),
                _lib.ptr(plan.inc),
                _lib.ptr(allb),
                stream,
            )
            self._cache[("bmask", "boundary")] = allb
        if region == "boundary":
            return allb
        if region not in FACES:
            raise ValueError(f"unknown region {region!r}; use 'boundary' or one of {sorted(FACES)}")
        if self.box is None:
            raise ValueError(f"region {region!r} needs a structured box mesh")
        axis, end = FACES[region]
        value = float(self.box[end][axis])
        out = torch.empty_like(allb)
        _lib.call("jno_mask_plane", _lib.ptr(self.points), self.n_nodes, axis, value, _lib.ptr(allb), _lib.ptr(out), stream)
        self._cache[key] = out
        return out


def kuhn_box(nx: int, ny: Optional[int] = None, nz: Optional[int] = None, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             device="cuda") -> Mesh:
    """Structured box, six Kuhn tets per voxel, generated on the device."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    N = (nx + 1) * (ny + 1) * (nz + 1)
    T = 6 * nx * ny * nz
    dev = torch.device(device)
    points = torch.empty((N, 3), dtype=torch.float64, device=dev)
    cells = torch.empty((T, 4), dtype=torch.int32, device=dev)
    lo_c = (C.c_double * 3)(*map(float, lo))
    hi_c = (C.c_double * 3)(*map(float, hi))
    _lib.call("jno_mesh_kuhn_box", nx, ny, nz, lo_c, hi_c, _lib.ptr(points), _lib.ptr(cells), _lib.stream_ptr(dev))
    return Mesh(points, cells, 1, (tuple(map(float, lo)), tuple(map(float, hi))), (nx, ny, nz))


class Shape:
    """``jno.Shape.box(x0, y0, z0, x1, y1, z1).structured(n).domain()`` for the tet lattice."""

    def __init__(self, lo, hi):
        self.lo, self.hi = tuple(map(float, lo)), tuple(map(float, hi))
        self._counts = None

    @staticmethod
    def box(x0=0.0, y0=0.0, z0=0.0, x1=1.0, y1=1.0, z1=1.0) -> "Shape":
        return Shape((min(x0, x1), min(y0, y1), min(z0, z1)), (max(x0, x1), max(y0, y1), max(z0, z1)))

    def structured(self, n) -> "Shape":
        counts = (n, n, n) if isinstance(n, int) else tuple(int(c) for c in n)
        if len(counts) != 3:
            raise ValueError("structured(n): give one count or three (nx, ny, nz)")
        # the reference clamps each axis to >= 2 cells (domain_class.py:2320-2324)
        self._counts = tuple(max(2, int(c)) for c in counts)
        return self

    def domain(self, device="cuda") -> Mesh:
        if self._counts is None:
            raise NotImplementedError("only .structured(n) lattices are meshed here (no gmsh path)")
        nx, ny, nz = self._counts
        return kuhn_box(nx, ny, nz, self.lo, self.hi, device)
