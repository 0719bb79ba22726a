"""Opaque plan handles of the C ABI (jno_plan_create / get / destroy): one plan per cell content,
reference counted, bit-identical to the oracle's pattern and to the Python layer's plan."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu


def _create(cells, n_nodes, vec=1):
    from jno_amd import _lib

    h = C.c_void_p()
    _lib.call("jno_plan_create", _lib.ptr(cells), cells.shape[0], cells.shape[1], vec, n_nodes, C.byref(h),
              _lib.stream_ptr(cells.device))
    return h


def _view(h):
    from jno_amd import _lib

    v = _lib.JnoPlanView()
    _lib.call("jno_plan_get", h, C.byref(v))
    return v


def _arr(ptr, n):
    # a device int32 array of the plan, copied to the host with the HIP runtime (kind 2: device to host)
    out = np.empty(n, dtype=np.int32)
    torch.cuda.synchronize()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(out.ctypes.data, ptr, 4 * n, 2) == 0
    return out


def _live():
    from jno_amd import _lib

    n = C.c_int64()
    _lib.call("jno_plan_live_count", C.byref(n))
    return int(n.value)


def test_plan_handles_share_one_plan_per_content(dev):
    import jno_amd as J
    from jno_amd import _lib

    base = _live()
    n = 5
    m = J.kuhn_box(n)
    h1 = _create(m.cells, m.n_nodes)
    v1 = _view(h1)
    # the pattern equals the oracle's (compress_plan's sorted unique pairs) and the Python plan
    p, t = O.kuhn_box(n)
    rp, cl, _, _ = O.pattern_struct(t, 1, p.shape[0])
    assert v1.n_rows == m.n_nodes and v1.nnz == len(cl) and v1.refs == 1
    assert np.array_equal(_arr(v1.rowptr, v1.n_rows + 1), rp)
    assert np.array_equal(_arr(v1.col, v1.nnz), cl)
    pl = m.plan()
    assert np.array_equal(_arr(v1.inc_ptr, m.n_nodes + 1), pl.inc_ptr.cpu().numpy())
    assert np.array_equal(_arr(v1.inc, 4 * m.n_cells), pl.inc.cpu().numpy())
    # the same content from another buffer: the same plan, one more reference
    h2 = _create(m.cells.clone(), m.n_nodes)
    assert h2.value == h1.value and _view(h1).refs == 2
    # other content (a relabelled mesh, and another size): other plans
    cells3 = m.cells.flip(0).contiguous()
    h3 = _create(cells3, m.n_nodes)
    m4 = J.kuhn_box(4)
    h4 = _create(m4.cells, m4.n_nodes)
    assert len({h1.value, h3.value, h4.value}) == 3
    assert _live() == base + 3
    # vec = 3 on the P2 cells: the vector pattern (node pattern times 3x3 blocks)
    from jno_amd.elasticity import promote_p2

    m4p2 = promote_p2(m4)
    h5 = _create(m4p2.cells, m4p2.n_nodes, vec=3)
    v5 = _view(h5)
    p2o, t2 = O.promote_p2(*O.kuhn_box(4))
    rp5, cl5, _, _ = O.pattern_struct(t2, 3, 3 * p2o.shape[0])
    assert v5.n_rows == 3 * m4p2.n_nodes and np.array_equal(_arr(v5.col, v5.nnz), cl5)
    assert np.array_equal(_arr(v5.rowptr, v5.n_rows + 1), rp5)
    # references drop one at a time; the last frees the plan, whose handle is then unknown
    _lib.call("jno_plan_destroy", h2)
    assert _view(h1).refs == 1
    for h in (h1, h3, h4, h5):
        _lib.call("jno_plan_destroy", h)
    assert _live() == base
    with pytest.raises(ValueError, match="not a live plan"):
        _lib.call("jno_plan_destroy", h1)
    with pytest.raises(ValueError, match="not a live plan"):
        _lib.call("jno_plan_get", h1, C.byref(_lib.JnoPlanView()))
    # a plan created again after the last reference went is a fresh build
    h6 = _create(m.cells, m.n_nodes)
    assert _view(h6).refs == 1 and np.array_equal(_arr(_view(h6).col, v1.nnz), cl)
    _lib.call("jno_plan_destroy", h6)


def test_plan_create_rejects_bad_arguments(dev):
    import jno_amd as J
    from jno_amd import _lib

    m = J.kuhn_box(2)
    h = C.c_void_p()
    for nloc, vec in ((0, 1), (11, 1), (4, 0), (4, 4)):
        with pytest.raises(ValueError):
            _lib.call("jno_plan_create", _lib.ptr(m.cells), m.n_cells, nloc, vec, m.n_nodes, C.byref(h),
                      _lib.stream_ptr(m.device))
    with pytest.raises(ValueError, match="empty"):
        _lib.call("jno_plan_create", _lib.ptr(m.cells), 0, 4, 1, m.n_nodes, C.byref(h), _lib.stream_ptr(m.device))
