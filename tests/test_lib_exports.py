"""The C-ABI library loads and exports every entry point include/jno_fem.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jno_fem.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jno_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for required in ("jno_mesh_kuhn_box", "jno_pattern_count", "jno_pattern_fill", "jno_assemble_p1",
                     "jno_dirichlet_eliminate", "jno_pcg_f64", "jno_pcg_mixed", "jno_adjoint_grad_k_p1"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from jno_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and every ctypes signature the host layer binds is a declared entry point
    assert set(_lib.SIGNATURES) <= set(declared_functions())


def test_library_version_and_error_channel():
    from jno_amd import _lib

    lib = _lib.load()
    assert b"gfx950" in lib.jno_version()
    # an argument error is reported without touching a device
    rc = lib.jno_mesh_kuhn_box(0, 1, 1, None, None, None, None, None)
    assert rc == _lib.JNO_EINVAL
    assert "cell counts" in _lib.last_error()
    with pytest.raises(ValueError, match="cell counts"):
        _lib.check(rc, "jno_mesh_kuhn_box")


def test_workspace_queries_are_pure():
    from jno_amd import _lib

    lib = _lib.load()
    assert lib.jno_pcg_workspace_size(1000, 0) > 1000 * 8 * 4
    assert lib.jno_pcg_workspace_size(1000, 1) > 1000 * 4 * 4
    assert lib.jno_pattern_workspace_size(6000, 4, 1331) > 1331 * 4


def test_missing_library_fails_loudly(tmp_path):
    from jno_amd import _lib

    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(ImportError, match="no CPU fallback"):
            _lib.load(tmp_path / "libjno_fem.so")
    finally:
        _lib._lib = saved


def test_library_is_built_for_gfx950():
    from jno_amd import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert isinstance(ctypes.CDLL(str(_lib.LIB_PATH)), ctypes.CDLL)
