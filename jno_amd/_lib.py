"""ctypes binding of libjno_fem.so (the C-ABI declared in include/jno_fem.h).

The HIP library is the only compute path: if it is missing or cannot be loaded, every entry point
raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libjno_fem.so"

JNO_OK = 0
JNO_EINVAL = -1
JNO_EHIP = -2
JNO_ENOTCONV = -3
JNO_ENONFINITE = -4
JNO_EWORKSPACE = -5
JNO_ECAPACITY = -6

JNO_FMT_CSR = 0
JNO_FMT_DIA = 1
JNO_FMT_BSR3 = 2
JNO_FMT_EBE = 3
JNO_DIA_MAX = 16


class JnoLaunchShape(C.Structure):
    _fields_ = [("grid_cap", C.c_int32), ("k1_vector", C.c_int32), ("block", C.c_int32), ("split", C.c_int32)]


class JnoTuning(C.Structure):
    _fields_ = [("iter", JnoLaunchShape * 2), ("chunk", C.c_int32), ("pacing", C.c_int32)]


class JnoEbe(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64),
        ("n_nodes", C.c_int64),
        ("n_tiles", C.c_int64),
        ("n_classes", C.c_int32),
        ("pad", C.c_int32),
        ("cells", C.c_void_p),
        ("order", C.c_void_p),
        ("tile_class", C.c_void_p),
        ("slot_mask", C.c_void_p),
        ("table64", C.c_void_p),
        ("table32", C.c_void_p),
        ("scale64", C.c_void_p),
        ("scale32", C.c_void_p),
        ("inc_ptr", C.c_void_p),
        ("inc", C.c_void_p),
        ("qpos", C.c_void_p),
        ("dmask", C.c_void_p),
        ("scratch", C.c_void_p),
        ("box_n", C.c_int32 * 3),
        ("box_class", C.c_int32 * 6),
        ("box_pad", C.c_int32),
        ("box_perm", C.c_void_p),
        ("box_mask", C.c_void_p),
        ("box_e64", C.c_void_p),
        ("box_e32", C.c_void_p),
        ("box_rows32", C.c_void_p),
    ]


class JnoOp(C.Structure):
    _fields_ = [
        ("format", C.c_int32),
        ("lanes", C.c_int32),
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("rowptr", C.c_void_p),
        ("col", C.c_void_p),
        ("vals64", C.c_void_p),
        ("vals32", C.c_void_p),
        ("ndiag", C.c_int32),
        ("offsets", C.c_int32 * JNO_DIA_MAX),
        ("tuning", C.c_void_p),  # const jno_tuning* (NULL: the library defaults)
        ("ebe", C.c_void_p),  # const jno_ebe* (JNO_FMT_EBE)
    ]


class JnoPcgStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("refinements", C.c_int64),
        ("relres", C.c_double),
        ("relres_recursive", C.c_double),
        ("seconds", C.c_double),
    ]


class JnoBatchItem(C.Structure):
    _fields_ = [("op", JnoOp), ("dinv", C.c_void_p), ("b", C.c_void_p), ("x", C.c_void_p)]


class JnoPlanView(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64),
        ("nnz", C.c_int64),
        ("n_cells", C.c_int64),
        ("n_nodes", C.c_int64),
        ("nloc", C.c_int32),
        ("vec", C.c_int32),
        ("max_row", C.c_int32),
        ("refs", C.c_int32),
        ("rowptr", C.c_void_p),
        ("col", C.c_void_p),
        ("inc_ptr", C.c_void_p),
        ("inc", C.c_void_p),
        ("key", C.c_uint64 * 2),
    ]


class JnoBatchTuning(C.Structure):
    _fields_ = [("one_kernel", C.c_int32), ("rows_per_entry", C.c_int32), ("grid", C.c_int32), ("ran_one_kernel", C.c_int32)]


class JnoError(RuntimeError):
    """A failed call into libjno_fem (carries the status code)."""

    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
D = C.c_double
SZ = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "jno_last_error": (C.c_int, [C.c_char_p, SZ]),
    "jno_version": (C.c_char_p, []),
    "jno_mesh_kuhn_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(D), C.POINTER(D), P, P, P]),
    "jno_pattern_workspace_size": (SZ, [I64, C.c_int, I64]),
    "jno_pattern_count": (C.c_int, [P, I64, C.c_int, C.c_int, I64, P, P, P, C.POINTER(I64), C.POINTER(I32), P, SZ, P]),
    "jno_pattern_fill": (C.c_int, [P, I64, C.c_int, C.c_int, I64, P, P, P, P, P]),
    "jno_hash_i32": (C.c_int, [P, I64, C.POINTER(C.c_uint64), P, SZ, P]),
    "jno_equal_i32": (C.c_int, [P, P, I64, C.POINTER(C.c_int), P, SZ, P]),
    "jno_boundary_mask": (C.c_int, [P, I64, C.c_int, I64, P, P, P, P]),
    "jno_mask_plane": (C.c_int, [P, I64, C.c_int, D, P, P, P]),
    "jno_assemble_p1": (C.c_int, [C.c_int, P, P, I64, I64, P, P, P, P, C.c_int, P, D, P, P, P, P, P, P, P]),
    "jno_assemble_p1_dia": (C.c_int, [C.c_int, P, P, I64, I64, P, P, P, P, C.c_int, P, D, P, P, P, P, P, P, C.c_int,
                                      C.POINTER(I32), P, P, C.POINTER(C.c_int), P, SZ, P]),
    "jno_pattern_kuhn_box_nnz": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(I64), C.POINTER(I32)]),
    "jno_pattern_kuhn_box": (C.c_int, [C.c_int, C.c_int, C.c_int, P, P, P, P, P]),
    "jno_mask_kuhn_box": (C.c_int, [C.c_int, C.c_int, C.c_int, P, P]),
    "jno_box_geometry": (C.c_int, [C.c_int, C.c_int, C.c_int, P, C.POINTER(D), P, SZ, P]),
    "jno_mesh_is_kuhn_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(D), C.POINTER(D), P, P,
                                       C.POINTER(C.c_int), P, SZ, P]),
    "jno_assemble_p1_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(D), P, P, D, P, P, P, P, P, P,
                                      C.c_int, C.POINTER(I32), P, P, P]),
    "jno_dirichlet_eliminate": (C.c_int, [I64, P, P, P, P, P, P, P]),
    "jno_jacobi": (C.c_int, [I64, P, P, P, P, P]),
    "jno_spmv_f64": (C.c_int, [I64, P, P, P, P, P, P]),
    "jno_drop_zeros_count": (C.c_int, [I64, P, P, P, C.POINTER(I64), P, SZ, P]),
    "jno_drop_zeros_fill": (C.c_int, [I64, P, P, P, P, P, P, P]),
    "jno_op_analyze_workspace_size": (SZ, []),
    "jno_op_analyze": (C.c_int, [I64, P, P, P, C.POINTER(I32), C.POINTER(I32), P, SZ, P]),
    "jno_op_fill_dia": (C.c_int, [I64, P, P, P, C.c_int, C.POINTER(I32), P, P, P]),
    "jno_op_fill_dia_checked": (C.c_int, [I64, P, P, P, C.c_int, C.POINTER(I32), P, P, C.POINTER(C.c_int), P, SZ, P]),
    "jno_op_bsr3_workspace_size": (SZ, []),
    "jno_op_bsr3_count": (C.c_int, [I64, P, P, C.POINTER(I64), P, SZ, P]),
    "jno_op_fill_bsr3": (C.c_int, [I64, P, P, P, P, P, P, P, P]),
    "jno_cast_f32": (C.c_int, [I64, P, P, P]),
    "jno_p2_cell_jacobians": (C.c_int, [P, P, I64, P, P]),
    "jno_ebe_class_blocks": (C.c_int, [P, P, C.c_int, P, D, P, P, P]),
    "jno_ebe_slots": (C.c_int, [I64, P, P, P, P, P, P]),
    "jno_ebe_scales": (C.c_int, [I64, P, P, P, P, P]),
    "jno_ebe_incidence": (C.c_int, [I64, P, P, I64, P, P, P]),
    "jno_ebe_diagonal": (C.c_int, [C.POINTER(JnoOp), C.c_int, P, P]),
    "jno_ebe_dirichlet_rhs": (C.c_int, [I64, P, P, P]),
    "jno_op_spmv_f64": (C.c_int, [C.POINTER(JnoOp), P, P, P]),
    "jno_op_spmv_f32": (C.c_int, [C.POINTER(JnoOp), P, P, P]),
    "jno_pcg_workspace_size": (SZ, [I64, C.c_int]),
    "jno_pcg_workspace_size_op": (SZ, [C.POINTER(JnoOp), C.c_int]),
    "jno_pcg_set_chunk": (C.c_int, [C.c_int]),
    "jno_pcg_set_pacing": (C.c_int, [C.c_int]),
    "jno_pcg_set_tuning": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "jno_pcg_get_tuning": (C.c_int, [C.c_int, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "jno_pcg_default_tuning": (C.c_int, [C.POINTER(JnoTuning)]),
    "jno_pcg_check_tuning": (C.c_int, [C.POINTER(JnoTuning)]),
    "jno_pcg_profile": (C.c_int, [C.c_int]),
    "jno_pcg_profile_read": (C.c_int, [C.POINTER(D), C.POINTER(I64)]),
    "jno_pcg_f64": (C.c_int, [C.POINTER(JnoOp), P, P, P, C.c_int, D, D, I64, C.POINTER(JnoPcgStats), P, SZ, P]),
    "jno_pcg_mixed": (
        C.c_int,
        [C.POINTER(JnoOp), P, P, P, C.c_int, D, D, I64, D, C.POINTER(JnoPcgStats), P, SZ, P],
    ),
    "jno_pcg_mixed_noalloc": (
        C.c_int,
        [C.POINTER(JnoOp), P, P, P, C.c_int, D, D, I64, D, C.POINTER(JnoPcgStats), P],
    ),
    "jno_adjoint_grad_k_p1": (C.c_int, [P, P, I64, P, P, P, P, P]),
    "jno_plan_create": (C.c_int, [P, I64, C.c_int, C.c_int, I64, C.POINTER(C.c_void_p), P]),
    "jno_plan_get": (C.c_int, [C.c_void_p, C.POINTER(JnoPlanView)]),
    "jno_plan_destroy": (C.c_int, [C.c_void_p]),
    "jno_plan_live_count": (C.c_int, [C.POINTER(I64)]),
    "jno_batch_workspace_size": (SZ, [C.c_int, C.POINTER(JnoBatchItem)]),
    "jno_batch_pcg_mixed": (
        C.c_int,
        [C.c_int, C.POINTER(JnoBatchItem), D, D, I64, D, C.c_int, C.POINTER(JnoPcgStats), P, SZ, P],
    ),
    "jno_batch_pcg_mixed_tuned": (
        C.c_int,
        [C.c_int, C.POINTER(JnoBatchItem), D, D, I64, D, C.c_int, C.POINTER(JnoBatchTuning), C.POINTER(JnoPcgStats), P,
         SZ, P],
    ),
    "jno_promote_p2_workspace_size": (SZ, [I64]),
    "jno_promote_p2": (C.c_int, [P, P, I64, I64, P, P, P, P, C.POINTER(I64), P, SZ, P]),
    "jno_color_workspace_size": (SZ, [I64]),
    "jno_color_cells": (C.c_int, [P, I64, C.c_int, P, P, P, C.POINTER(I32), I32, C.POINTER(I32), P, SZ, P]),
    "jno_assemble_p2_elasticity": (C.c_int, [P, P, I64, P, P, I64, P, D, P, C.POINTER(I32), C.c_int, P, P, P]),
    "jno_p2_pair_rank": (C.c_int, [P, I64, P, P, P, C.POINTER(I32), P, C.c_size_t, P]),
    "jno_assemble_p2_classes": (C.c_int, [I64, P, P, P, P, P, P, P, P, P, C.c_int, P, P]),
    "jno_traction_p2": (C.c_int, [P, P, I64, P, P, C.c_int, D, C.POINTER(D), P, P]),
    "jno_compliance_grad_p2": (C.c_int, [P, P, I64, D, P, P, P]),
    "jno_theta_operators": (C.c_int, [I64, P, P, P, P, P, D, D, P, P, P, P]),
    "jno_theta_operators_dia": (C.c_int, [I64, C.c_int, C.POINTER(I32), P, C.c_int, C.POINTER(I32), P, P, D, D, P, P,
                                          P, P, P]),
    "jno_theta_rhs": (C.c_int, [C.POINTER(JnoOp), P, P, P, P, D, D, P, P]),
    "jno_theta_mix": (C.c_int, [I64, P, P, P, P, D, P, P]),
}

_lib = None


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libjno_fem.so (built by ``python -m jno_amd.build``); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is missing: the jno_amd HIP library is not built. Run `python -m jno_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback."
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(2048)
    lib.jno_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


def check(status: int, what: str = "") -> int:
    if status == JNO_OK:
        return status
    msg = last_error() or f"status {status}"
    if status == JNO_EINVAL:
        raise ValueError(f"{what}: {msg}" if what else msg)
    raise JnoError(status, f"{what}: {msg}" if what else msg)


def call(name: str, *args) -> int:
    return check(getattr(load(), name)(*args), name)


def ptr(t) -> C.c_void_p | None:
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stream_ptr(device=None) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu(t: torch.Tensor, name: str):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (got {t.device}); jno_amd has no CPU path")
