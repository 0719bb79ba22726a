"""Element-by-element (matrix-free) operator of the P2 elasticity form.

The analogue of the reference's ``ElementOperator`` (jno/utils/solver/element_operator.py:1-178):
the assembled operator kept as per-cell dense blocks plus the connectivity instead of a global
sparse matrix, with the ``LinearOperator`` subset the Krylov solvers use -- ``mv`` / ``@`` /
``__call__``, ``shape``, ``T``, ``diag()`` -- ``bcoo = None``, and ``assemble()`` to build the
matrix only when asked.  Here a cell's block is ``E_e * K_c``: cells whose vertex differences are
bitwise equal (the 6 tet shapes of a structured Kuhn box) share the block ``K_c`` of their geometry
class, so the operator stores a class table, one scale per cell and the connectivity -- 48 bytes
per cell instead of the 900 values a block would hold, and no per-nonzero indices.  The operator is
the symmetrically eliminated one the CSR path solves (``y = P K P x + (I - P) x``).  Its SpMV runs
the element products on MFMA and an atomic-free node gather (``csrc/ebe.hip``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

TILE = 16
PITCH = 32


class EbeGeometry:
    """The mesh-, form- and constraint-dependent part of the operator (built once per system):
    geometry classes, slot order (cells grouped by class in tiles of 16), class blocks, the
    constrained-dof bits per slot and the node incidence in slot numbering."""

    def __init__(self, system):
        m, plan = system.mesh, system.plan
        dev = m.device
        T, N = m.n_cells, m.n_nodes
        lib = _lib.load()
        stream = _lib.stream_ptr(dev)
        jac = torch.empty(max(9 * T, 1), dtype=torch.float64, device=dev)
        _lib.call("jno_p2_cell_jacobians", _lib.ptr(m.points), _lib.ptr(m.cells), T, _lib.ptr(jac), stream)
        # classes: cells with bitwise-equal vertex differences (compared as integers, so -0.0 and 0.0
        # stay apart); found on the host as the reference finds its pattern (np.unique)
        J = jac[: 9 * T].view(torch.int64).reshape(T, 9).cpu().numpy()
        _, first, cls = np.unique(J, axis=0, return_index=True, return_inverse=True)
        cls = cls.reshape(-1)
        n_classes = int(first.shape[0])
        # a tile holds one class: past about one class per tile of cells the padding outweighs the
        # saving (small meshes may keep up to 64 classes)
        if T and n_classes > max(64, T // TILE):
            raise ValueError(
                f"the element-by-element operator needs repeated cell geometry: {n_classes} classes for {T} cells "
                "(use the 'bsr3' or 'csr' layout for unstructured meshes)"
            )
        # slots: the cells of each class in ascending order, cut into tiles of 16 (the last tile of a
        # class padded with -1), and the tiles ordered by their first cell, so that consecutive tiles
        # cover neighbouring cells (the SpMV deals contiguous tile ranges to the XCDs)
        counts = np.bincount(cls, minlength=n_classes)
        tiles = (counts + TILE - 1) // TILE
        by = np.argsort(cls, kind="stable")
        cstart = np.concatenate([[0], np.cumsum(counts)])
        tstart = np.concatenate([[0], np.cumsum(tiles)])
        rank = np.arange(T) - cstart[cls[by]]
        ctile = tstart[cls[by]] + rank // TILE  # class-major tile of each cell (in `by` order)
        tile_first = np.full(int(tstart[-1]), T, np.int64)
        np.minimum.at(tile_first, ctile, by)
        tile_rank = np.empty(int(tstart[-1]), np.int64)
        tile_rank[np.argsort(tile_first, kind="stable")] = np.arange(int(tstart[-1]))
        slot_of = np.empty(T, np.int64)
        slot_of[by] = tile_rank[ctile] * TILE + rank % TILE
        n_slots = int(tstart[-1]) * TILE
        order = np.full(n_slots, -1, np.int32)
        order[slot_of] = np.arange(T, dtype=np.int32)
        tile_class = np.empty(int(tstart[-1]), np.int32)
        tile_class[tile_rank] = np.repeat(np.arange(n_classes, dtype=np.int32), tiles)
        self.n_cells, self.n_nodes, self.n_classes = T, N, n_classes
        self.n_tiles = int(tile_class.shape[0])
        self.order = torch.tensor(order, device=dev)
        self.cell_class = torch.tensor(cls.astype(np.int32), device=dev)
        self.tile_class = torch.tensor(tile_class, device=dev)
        slot_of_d = torch.tensor(slot_of.astype(np.int32), device=dev)
        rep = torch.tensor(first.astype(np.int32), device=dev)
        self.table64 = torch.empty(max(n_classes, 1) * PITCH * PITCH, dtype=torch.float64, device=dev)
        self.table32 = torch.empty(max(n_classes, 1) * PITCH * PITCH, dtype=torch.float32, device=dev)
        _lib.call("jno_ebe_class_blocks", _lib.ptr(m.points), _lib.ptr(m.cells), n_classes, _lib.ptr(rep), system.nu,
                  _lib.ptr(self.table64), _lib.ptr(self.table32), stream)
        self.inc_ptr = plan.inc_ptr
        self.inc = torch.empty_like(plan.inc)
        self.qpos = torch.empty(max(10 * n_slots, 1), dtype=torch.int32, device=dev)
        _lib.call("jno_ebe_incidence", int(plan.inc.shape[0]), _lib.ptr(plan.inc), _lib.ptr(slot_of_d), n_slots,
                  _lib.ptr(self.inc), _lib.ptr(self.qpos), stream)
        self.dmask = system.dmask if system.dmask is not None else torch.zeros(3 * N, dtype=torch.uint8, device=dev)
        self.box = self._box_map(m, cls, n_classes, stream)
        self.n_slots = n_slots
        self.slot_mask = torch.empty(max(n_slots, 1), dtype=torch.int32, device=dev)
        self.cells = torch.empty(max(10 * n_slots, 1), dtype=torch.int32, device=dev)
        _lib.call("jno_ebe_slots", n_slots, _lib.ptr(self.order), _lib.ptr(m.cells), _lib.ptr(self.dmask),
                  _lib.ptr(self.cells), _lib.ptr(self.slot_mask), stream)
        del lib

    def _box_map(self, m, cls, n_classes, stream):
        """The half-lattice data of the structured-box SpMV when the mesh is the P2 Kuhn box of its grid
        and every cell of one Kuhn type has one class (checked on the device), else None."""
        grid = getattr(m, "grid", None)
        T, N = m.n_cells, m.n_nodes
        if grid is None or T == 0:
            return None
        nx, ny, nz = grid
        if T != 6 * nx * ny * nz or N != (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1):
            return None
        per_type = cls.reshape(-1, 6)
        if not (per_type == per_type[0]).all():
            return None
        dev = m.device
        perm = torch.empty(N, dtype=torch.int32, device=dev)
        mlat = torch.empty(N, dtype=torch.uint8, device=dev)
        ws = torch.empty(N + 1, dtype=torch.int32, device=dev)
        ok = C.c_int32(0)
        _lib.call("jno_ebe_box_map", nx, ny, nz, _lib.ptr(m.cells), T, N, _lib.ptr(self.dmask), _lib.ptr(perm),
                  _lib.ptr(mlat), _lib.ptr(ws), C.byref(ok), stream)
        if not ok.value:
            return None
        cls6 = (C.c_int32 * 6)(*[int(c) for c in per_type[0]])
        rows = torch.empty(6 * 10 * 10 * 16, dtype=torch.float32, device=dev)
        _lib.call("jno_ebe_box_rows", _lib.ptr(self.table32), n_classes, cls6, _lib.ptr(rows), stream)
        return {"grid": (nx, ny, nz), "perm": perm, "mask": mlat, "rows": rows, "class": cls6}

    @property
    def nbytes(self) -> int:
        ts = (self.order, self.tile_class, self.table64, self.table32, self.inc, self.qpos, self.slot_mask, self.cells,
              self.cell_class)
        if self.box is not None:
            ts += (self.box["perm"], self.box["mask"], self.box["rows"])
        return int(sum(t.numel() * t.element_size() for t in ts))


class ElementOperator:
    """The eliminated P2 elasticity operator at per-cell moduli ``E``, element by element (see the
    module docstring).  ``solve.cg(...)(op, b, M=precond.jacobi())`` solves on it directly."""

    format = "ebe"

    def __init__(self, system, E: torch.Tensor, structured: bool = True):
        geo = getattr(system, "_ebe_geometry", None)
        if geo is None:
            geo = EbeGeometry(system)
            system._ebe_geometry = geo
        self.system, self.geometry = system, geo
        dev = system.mesh.device
        self.device = dev
        self.n = 3 * geo.n_nodes
        self.E = E.to(device=dev, dtype=torch.float64).contiguous()
        # per-slot scales (slot order, 0 for padding), FP64 and FP32
        self.scale64 = torch.empty(max(geo.n_slots, 1), dtype=torch.float64, device=dev)
        self.scale32 = torch.empty(max(geo.n_slots, 1), dtype=torch.float32, device=dev)
        _lib.call("jno_ebe_scales", geo.n_slots, _lib.ptr(geo.order), _lib.ptr(self.E), _lib.ptr(self.scale64),
                  _lib.ptr(self.scale32), _lib.stream_ptr(dev))
        self.scratch = torch.empty(max(30 * geo.n_cells, 1), dtype=torch.float64, device=dev)
        e = _lib.JnoEbe()
        e.n_cells, e.n_nodes, e.n_tiles, e.n_classes = geo.n_cells, geo.n_nodes, geo.n_tiles, geo.n_classes
        e.cells, e.order, e.tile_class = geo.cells.data_ptr(), geo.order.data_ptr(), geo.tile_class.data_ptr()
        e.slot_mask = geo.slot_mask.data_ptr()
        e.table64, e.table32 = geo.table64.data_ptr(), geo.table32.data_ptr()
        e.scale64, e.scale32 = self.scale64.data_ptr(), self.scale32.data_ptr()
        e.inc_ptr, e.inc, e.dmask = geo.inc_ptr.data_ptr(), geo.inc.data_ptr(), geo.dmask.data_ptr()
        e.qpos = geo.qpos.data_ptr()
        e.scratch = self.scratch.data_ptr()
        # the structured Kuhn box: FP32 products on the half lattice (ebe_box.hip), no product buffer
        self.structured = bool(structured and geo.box is not None)
        if self.structured:
            bx = geo.box
            self.e32 = torch.empty(max(geo.n_cells, 1), dtype=torch.float32, device=dev)
            _lib.call("jno_cast_f32", geo.n_cells, _lib.ptr(self.E), _lib.ptr(self.e32), _lib.stream_ptr(dev))
            for d in range(3):
                e.box_n[d] = bx["grid"][d]
            for t in range(6):
                e.box_class[t] = bx["class"][t]
            e.box_perm, e.box_mask, e.box_rows32 = bx["perm"].data_ptr(), bx["mask"].data_ptr(), bx["rows"].data_ptr()
            e.box_e64, e.box_e32 = self.E.data_ptr(), self.e32.data_ptr()
        self._ebe = e
        op = _lib.JnoOp()
        op.format = _lib.JNO_FMT_EBE
        op.lanes = 1
        op.n = self.n
        op.nnz = system.plan.nnz
        op.vals64, op.vals32 = geo.table64.data_ptr(), geo.table32.data_ptr()
        op.ebe = C.addressof(e)
        self.op = op
        self.tuning = None
        self._dinv = None
        self._cache: dict = {}

    # -- the LinearOperator subset of the reference (element_operator.py:89-167) ------------------
    @property
    def shape(self):
        return (self.n, self.n)

    @property
    def bcoo(self):
        """None: there is no assembled matrix (see :meth:`assemble`)."""
        return None

    @property
    def dtype(self):
        return torch.float64

    @property
    def nbytes(self) -> int:
        return self.geometry.nbytes + int(self.scale64.numel() * 12)

    def mv(self, x: torch.Tensor) -> torch.Tensor:
        stream = _lib.stream_ptr(self.device)
        if x.dtype == torch.float32:
            y = torch.empty(self.n, dtype=torch.float32, device=self.device)
            _lib.call("jno_op_spmv_f32", C.byref(self.op), _lib.ptr(x.contiguous()), _lib.ptr(y), stream)
            return y
        y = torch.empty(self.n, dtype=torch.float64, device=self.device)
        _lib.call("jno_op_spmv_f64", C.byref(self.op), _lib.ptr(x.to(torch.float64).contiguous()), _lib.ptr(y), stream)
        return y

    __matmul__ = mv
    __call__ = mv

    @property
    def T(self) -> "ElementOperator":
        """The operator is symmetric (each block is, and the elimination keeps it so)."""
        return self

    def diag_inv(self) -> torch.Tensor:
        """Jacobi inverse of the exact diagonal, |d| > 1e-30 guard (jno/precond.py:105-109)."""
        if self._dinv is None:
            d = torch.empty(self.n, dtype=torch.float64, device=self.device)
            _lib.call("jno_ebe_diagonal", C.byref(self.op), 1, _lib.ptr(d), _lib.stream_ptr(self.device))
            self._dinv = d
        return self._dinv

    def diag(self) -> torch.Tensor:
        """The exact diagonal (ElementOperator.diag, element_operator.py:150-157)."""
        d = torch.empty(self.n, dtype=torch.float64, device=self.device)
        _lib.call("jno_ebe_diagonal", C.byref(self.op), 0, _lib.ptr(d), _lib.stream_ptr(self.device))
        return d

    def assemble(self):
        """The assembled CSR of the same operator, built explicitly (never implicitly)."""
        return self.system.evaluate({self.system.param: self.E})[0]

    def dense(self) -> torch.Tensor:
        return self.assemble().dense()

    def describe(self) -> dict:
        g = self.geometry
        return {"format": "ebe", "classes": g.n_classes, "cells": g.n_cells, "tiles": g.n_tiles,
                "structured": self.structured}

    def __repr__(self) -> str:
        g = self.geometry
        return (f"ElementOperator(shape={self.shape}, classes={g.n_classes}, cells={g.n_cells}, "
                f"{self.nbytes / 2**20:.1f} MiB)")
