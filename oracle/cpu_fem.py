"""ctypes wrapper of oracle/libcpu_fem.so, the C + OpenMP FP64 restatement of the P1 path.

TEST INFRASTRUCTURE ONLY (timed CPU baseline of bench.py, second checker of the tests)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_fem.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        lib.oracle_p1_pattern_count.restype = C.c_int64
        lib.oracle_p1_pattern_count.argtypes = [P, C.c_int64, C.c_int64, P, P, P]
        lib.oracle_p1_pattern_fill.restype = None
        lib.oracle_p1_pattern_fill.argtypes = [P, C.c_int64, P, P, P, P]
        lib.oracle_p1_assemble.restype = None
        lib.oracle_p1_assemble.argtypes = [P, P, C.c_int64, P, P, P, P, P, C.c_double, P, P, P, P, P]
        lib.oracle_pcg.restype = C.c_int64
        lib.oracle_pcg.argtypes = [C.c_int64, P, P, P, P, P, P, C.c_double, C.c_double, C.c_int64,
                                   C.POINTER(C.c_double)]
        lib.oracle_num_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class P1System:
    """Pattern + assembled, eliminated P1 Poisson system on the CPU."""

    def __init__(self, points, cells, k, f=1.0, dmask=None, g=None):
        lib = load()
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        T = self.cells.shape[0]
        N = self.points.shape[0]
        self.n = N
        t0 = time.perf_counter()
        self.inc_ptr = np.empty(N + 1, np.int32)
        self.inc = np.empty(4 * T, np.int32)
        self.rowptr = np.empty(N + 1, np.int32)
        nnz = lib.oracle_p1_pattern_count(_p(self.cells), T, N, _p(self.inc_ptr), _p(self.inc), _p(self.rowptr))
        self.col = np.empty(nnz, np.int32)
        lib.oracle_p1_pattern_fill(_p(self.cells), N, _p(self.inc_ptr), _p(self.inc), _p(self.rowptr), _p(self.col))
        t1 = time.perf_counter()
        self.vals = np.empty(nnz, np.float64)
        self.b = np.empty(N, np.float64)
        self.dinv = np.empty(N, np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        dm = None if dmask is None else np.ascontiguousarray(dmask, dtype=np.uint8)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        self._args = (k, float(f), dm, gg)
        self.assemble()
        t2 = time.perf_counter()
        self.t_pattern, self.t_assemble = t1 - t0, t2 - t1

    def assemble(self) -> float:
        """Re-assemble A, b, dinv on the plan (the per-step work); returns the seconds it took."""
        k, f, dm, gg = self._args
        t0 = time.perf_counter()
        load().oracle_p1_assemble(_p(self.points), _p(self.cells), self.n, _p(self.inc_ptr), _p(self.inc),
                                  _p(self.rowptr), _p(self.col), _p(k), f, _p(dm), _p(gg), _p(self.vals), _p(self.b),
                                  _p(self.dinv))
        return time.perf_counter() - t0

    def pcg(self, tol=1e-10, atol=0.0, maxiter=20000):
        x = np.empty(self.n, np.float64)
        rel = C.c_double(0.0)
        t0 = time.perf_counter()
        it = load().oracle_pcg(self.n, _p(self.rowptr), _p(self.col), _p(self.vals), _p(self.dinv), _p(self.b), _p(x),
                               tol, atol, maxiter, C.byref(rel))
        return x, int(it), float(rel.value), time.perf_counter() - t0


def threads() -> int:
    return int(load().oracle_num_threads())
