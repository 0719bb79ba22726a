"""Solver operators: the device layout the PCG kernels stream, derived from an assembled CSR.

Three layouts, picked per operator (``format="auto"`` lets :mod:`jno_amd.autotune` time them):
  * ``dia``  symmetric diagonal storage -- the main and upper diagonals of the nonzero pattern,
             the lower ones read through symmetry.  Index-free and fully coalesced; it applies when
             the nonzero entries lie on at most 31 distinct diagonals and the matrix is bitwise
             symmetric (structured meshes, e.g. the Kuhn lattice).
  * ``bsr3`` 3x3-block CSR for node-interleaved vector fields (vec = 3: P2 elasticity): one
             column index per 9 values, each block row's entries stored entry-major so a lane per
             block reads them coalesced; ``lanes`` lanes per block row.  The stored pattern is the
             structural one (zeros inside a block are kept).
  * ``csr``  CSR with ``lanes`` lanes of a wavefront per row (1..64), exact zeros dropped first
             (``compress_eager``, jno/utils/solver/fem_utils.py:4063-4091).
Both carry FP64 values (plain PCG, FP64 residuals) and FP32 values (mixed-precision inner loop).
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import torch

from . import _lib
from .fem import CSR


# DIA offsets of earlier operators, per pattern (the index tensors of a plan), fewest diagonals
# first: a re-assembled operator on the same plan reuses them through the checked fill instead of a
# new analysis.  One pattern can carry several sets (the stiffness has exact zeros on diagonals
# where the mass or a transient step operator of the same plan does not).  Entries are keyed by the
# identity of the live index tensors and dropped when those are freed, so a later pattern at a
# reused address never sees them.
_DIA_OFFSETS: dict = {}


def _pattern_key(A: CSR):
    return (id(A.rowptr), id(A.col), A.shape[0], A.nnz, A.family)


def _forget(key):
    _DIA_OFFSETS.pop(key, None)


def _remember_offsets(A: CSR, nd, offs, chosen):
    key = _pattern_key(A)
    if key not in _DIA_OFFSETS:
        weakref.finalize(A.rowptr, _forget, key)
        weakref.finalize(A.col, _forget, key)
    sets = _DIA_OFFSETS.setdefault(key, [])
    for i, (nd0, offs0, chosen0) in enumerate(sets):
        if offs0 == offs:
            sets[i] = (nd0, offs0, chosen0 or chosen)
            return
    sets.append((nd, offs, chosen))
    sets.sort(key=lambda e: e[0])


def chosen_dia_offsets(A: CSR):
    """(ndiag, offsets) with the fewest diagonals that the format choice picked for an earlier
    operator on A's pattern, or None.  The assembly uses them to write the DIA layout in the same
    pass (and falls back when A has entries off them)."""
    for nd, offs, chosen in _DIA_OFFSETS.get(_pattern_key(A), []):
        if chosen:
            return nd, offs
    return None


class SolverOperator:
    def __init__(self, A: CSR, fmt: str = "auto", lanes: Optional[int] = None, fp32: bool = True):
        self.A = A
        self.n = A.shape[0]
        self.device = A.device
        self._keep = []  # tensors referenced by the C struct
        self.op = _lib.JnoOp()
        self.tuning = None  # a jno_amd.solve.Tuning attached by the autotuner (None: library defaults)
        fmt = fmt.lower()
        if fmt not in ("auto", "dia", "csr", "bsr3"):
            raise ValueError("format must be 'auto', 'dia', 'bsr3' or 'csr'")
        self.ndiag = -1
        sets = _DIA_OFFSETS.get(_pattern_key(A), []) if fmt in ("auto", "dia") else []
        # under "auto", only offsets that the format choice itself picked are reused
        for nd, offs, chosen in sets:
            if (fmt == "dia" or chosen) and self._build_dia_checked(nd, offs, fp32):
                self.format = "dia"
                return
        chosen = fmt == "auto"
        if fmt == "auto":
            from .autotune import choose_format

            fmt, lanes = choose_format(A)
        if fmt == "dia":
            nd, offs = self.analyze(A)
            if nd < 0:
                raise ValueError("this operator has no symmetric DIA layout (too many diagonals or not symmetric)")
            self._build_dia(nd, offs, fp32)
            _remember_offsets(A, nd, offs, chosen)
        elif fmt == "bsr3":
            nb = self.bsr3_blocks(A)
            if nb < 0:
                raise ValueError("this operator has no BSR3 layout (not a node pattern times full 3x3 blocks)")
            self._build_bsr3(nb, lanes or 16, fp32)
        else:
            self._build_csr(A.drop_zeros(), lanes or 4, fp32)
        self.format = fmt

    @classmethod
    def from_dia(cls, A: CSR, nd, offs, d64, d32):
        """Operator on DIA buffers that were filled elsewhere (the fused assembly)."""
        op = cls.__new__(cls)
        op.A = A
        op.n = A.shape[0]
        op.device = A.device
        op._keep = []
        op.op = _lib.JnoOp()
        op.tuning = None
        op._set_dia(nd, list(offs), d64, d32)
        op.format = "dia"
        return op

    @staticmethod
    def analyze(A: CSR):
        """(ndiag, offsets) of the nonzero pattern, ndiag = -1 when DIA does not apply."""
        lib = _lib.load()
        wsb = lib.jno_op_analyze_workspace_size()
        ws = torch.empty(wsb, dtype=torch.uint8, device=A.device)
        nd = C.c_int32(0)
        offs = (C.c_int32 * _lib.JNO_DIA_MAX)()
        _lib.call("jno_op_analyze", A.shape[0], _lib.ptr(A.rowptr), _lib.ptr(A.col), _lib.ptr(A.vals), C.byref(nd),
                  offs, _lib.ptr(ws), wsb, _lib.stream_ptr(A.device))
        return int(nd.value), [int(offs[i]) for i in range(max(nd.value, 0))]

    @staticmethod
    def bsr3_blocks(A: CSR) -> int:
        """Number of 3x3 blocks of the BSR3 layout of ``A``, -1 when it does not apply."""
        lib = _lib.load()
        wsb = lib.jno_op_bsr3_workspace_size()
        ws = torch.empty(wsb, dtype=torch.uint8, device=A.device)
        nb = C.c_int64(-1)
        _lib.call("jno_op_bsr3_count", A.shape[0], _lib.ptr(A.rowptr), _lib.ptr(A.col), C.byref(nb), _lib.ptr(ws), wsb,
                  _lib.stream_ptr(A.device))
        return int(nb.value)

    def _build_bsr3(self, nb, lanes, fp32):
        if lanes not in (1, 2, 4, 8, 16, 32, 64):
            raise ValueError("lanes per block row must be a power of two in 1..64")
        n, A = self.n, self.A
        browptr = torch.empty(n // 3 + 1, dtype=torch.int32, device=self.device)
        bcol = torch.empty(max(nb, 1), dtype=torch.int32, device=self.device)
        v64 = torch.empty(max(9 * nb, 1), dtype=torch.float64, device=self.device)
        v32 = torch.empty(max(9 * nb, 1), dtype=torch.float32, device=self.device) if fp32 else None
        _lib.call("jno_op_fill_bsr3", n, _lib.ptr(A.rowptr), _lib.ptr(A.col), _lib.ptr(A.vals), _lib.ptr(browptr),
                  _lib.ptr(bcol), _lib.ptr(v64), _lib.ptr(v32), _lib.stream_ptr(self.device))
        self._keep += [browptr, bcol, v64, v32]
        op = self.op
        op.format = _lib.JNO_FMT_BSR3
        op.lanes = lanes
        op.n = n
        op.nnz = 9 * nb
        op.rowptr = browptr.data_ptr()
        op.col = bcol.data_ptr()
        op.vals64 = v64.data_ptr()
        op.vals32 = v32.data_ptr() if v32 is not None else None
        op.ndiag = 0
        self.lanes = lanes
        self.nnz_stored = 9 * nb

    def _build_dia_checked(self, nd, offs, fp32) -> bool:
        """DIA layout with earlier offsets, verified on the device to reproduce A; False (nothing kept)
        when it does not."""
        n, A = self.n, self.A
        ld = (n + 63) // 64 * 64
        d64 = torch.empty(nd * ld, dtype=torch.float64, device=self.device)
        d32 = torch.empty(nd * ld, dtype=torch.float32, device=self.device) if fp32 else None
        ws = torch.empty(256, dtype=torch.uint8, device=self.device)
        offs_c = (C.c_int32 * _lib.JNO_DIA_MAX)(*offs)
        ok = C.c_int(0)
        _lib.call("jno_op_fill_dia_checked", n, _lib.ptr(A.rowptr), _lib.ptr(A.col), _lib.ptr(A.vals), nd, offs_c,
                  _lib.ptr(d64), _lib.ptr(d32), C.byref(ok), _lib.ptr(ws), 256, _lib.stream_ptr(self.device))
        if not ok.value:
            return False
        self._set_dia(nd, offs, d64, d32)
        return True

    def _build_dia(self, nd, offs, fp32):
        n, A = self.n, self.A
        ld = (n + 63) // 64 * 64  # diagonals padded to 64 rows (aligned vector loads)
        d64 = torch.empty(nd * ld, dtype=torch.float64, device=self.device)
        d32 = torch.empty(nd * ld, dtype=torch.float32, device=self.device) if fp32 else None
        offs_c = (C.c_int32 * _lib.JNO_DIA_MAX)(*offs)
        _lib.call("jno_op_fill_dia", n, _lib.ptr(A.rowptr), _lib.ptr(A.col), _lib.ptr(A.vals), nd, offs_c, _lib.ptr(d64),
                  _lib.ptr(d32), _lib.stream_ptr(self.device))
        self._set_dia(nd, offs, d64, d32)

    def _set_dia(self, nd, offs, d64, d32):
        n, A = self.n, self.A
        self._keep += [d64, d32]
        op = self.op
        op.format = _lib.JNO_FMT_DIA
        op.lanes = 1
        op.n = n
        op.nnz = A.nnz
        op.vals64 = d64.data_ptr()
        op.vals32 = d32.data_ptr() if d32 is not None else None
        op.ndiag = nd
        for i, o in enumerate(offs):
            op.offsets[i] = o
        self.ndiag = nd
        self.offsets = offs
        self.lanes = 1
        self.nnz_stored = A.nnz

    def _build_csr(self, Az: CSR, lanes, fp32):
        if lanes not in (1, 2, 4, 8, 16, 32, 64):
            raise ValueError("lanes per row must be a power of two in 1..64")
        v32 = None
        if fp32:
            v32 = torch.empty(max(Az.nnz, 1), dtype=torch.float32, device=self.device)
            _lib.call("jno_cast_f32", Az.nnz, _lib.ptr(Az.vals), _lib.ptr(v32), _lib.stream_ptr(self.device))
        self._keep += [Az.rowptr, Az.col, Az.vals, v32]
        op = self.op
        op.format = _lib.JNO_FMT_CSR
        op.lanes = lanes
        op.n = self.n
        op.nnz = Az.nnz
        op.rowptr = Az.rowptr.data_ptr()
        op.col = Az.col.data_ptr()
        op.vals64 = Az.vals.data_ptr()
        op.vals32 = v32.data_ptr() if v32 is not None else None
        op.ndiag = 0
        self.lanes = lanes
        self.nnz_stored = Az.nnz

    def mv(self, x: torch.Tensor) -> torch.Tensor:
        if x.dtype == torch.float32:
            y = torch.empty(self.n, dtype=torch.float32, device=self.device)
            _lib.call("jno_op_spmv_f32", C.byref(self.op), _lib.ptr(x.contiguous()), _lib.ptr(y), _lib.stream_ptr(self.device))
            return y
        y = torch.empty(self.n, dtype=torch.float64, device=self.device)
        _lib.call("jno_op_spmv_f64", C.byref(self.op), _lib.ptr(x.to(torch.float64).contiguous()), _lib.ptr(y),
                  _lib.stream_ptr(self.device))
        return y

    def describe(self) -> dict:
        if self.format == "dia":
            return {"format": "dia", "ndiag": self.ndiag, "offsets": self.offsets}
        return {"format": self.format, "lanes": self.lanes, "nnz": self.nnz_stored}
