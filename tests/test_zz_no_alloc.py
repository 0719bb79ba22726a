"""The allocation-free mixed PCG entry point (jno_pcg_mixed_noalloc) against jno_pcg_mixed on the
same operators: calls in a row on one GPU, each smaller than the one before it, give bit-identical
solutions and statistics (the static scratch keeps the previous call's contents, so this also
checks that everything that must start at zero is cleared); an operator beyond the scratch is
rejected before any work is queued."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _both(A, b, op, tol=1e-10, rf=0.1):
    from jno_amd import _lib

    lib = _lib.load()
    n = A.shape[0]
    dinv = A.diag_inv()
    stream = _lib.stream_ptr(A.device)
    wsb = lib.jno_pcg_workspace_size_op(C.byref(op.op), 1)
    ws = torch.empty(wsb, dtype=torch.uint8, device=A.device)
    x1 = torch.empty(n, dtype=torch.float64, device=A.device)
    s1 = _lib.JnoPcgStats()
    rc1 = lib.jno_pcg_mixed(C.byref(op.op), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x1), 0, tol, 0.0, 20000, rf,
                            C.byref(s1), _lib.ptr(ws), wsb, stream)
    x2 = torch.full((n,), float("nan"), dtype=torch.float64, device=A.device)
    s2 = _lib.JnoPcgStats()
    rc2 = lib.jno_pcg_mixed_noalloc(C.byref(op.op), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x2), 0, tol, 0.0, 20000, rf,
                                    C.byref(s2), stream)
    return rc1, s1, _np(x1), rc2, s2, _np(x2)


def _same(r):
    rc1, s1, x1, rc2, s2, x2 = r
    assert rc1 == 0 and rc2 == 0, (rc1, rc2)
    assert np.array_equal(x1, x2)
    assert (s1.iterations, s1.refinements) == (s2.iterations, s2.refinements)
    assert s1.relres == s2.relres and s2.relres <= 1e-10


def test_noalloc_matches_the_workspace_entry_point(dev):
    import jno_amd as J
    from jno_amd import solve
    from jno_amd.elasticity import ElasticitySystem

    # P1 Poisson on DIA layouts (the one-kernel FP32 iteration), 40^3 then 16^3 rows
    for n in (40, 16):
        k = O.k_lognormal(6 * n**3, 70 + n)
        A, b = J.fem_poisson(J.kuhn_box(n)).evaluate({"k": torch.tensor(k, device="cuda")})
        op = solve.cg(precision="mixed").operator(A)
        assert op.format == "dia"
        r = _both(A, b, op)
        _same(r)
        if n == 16:
            p, t = O.kuhn_box(n)
            Ao, bo, _ = O.poisson_p1_system(p, t, k)
            xd = O.direct_solve(Ao, bo)
            assert np.linalg.norm(r[5] - xd) / np.linalg.norm(xd) <= 1e-9
    # P2 elasticity on BSR3 and CSR layouts, after the larger DIA calls left the scratch dirty
    E3 = np.exp(0.3 * np.random.default_rng(8).standard_normal(6 * 27))
    s3 = ElasticitySystem(J.kuhn_box(3))
    A3, b3 = s3.evaluate({"E": torch.tensor(E3, device="cuda")})
    for fmt in ("bsr3", "csr"):
        _same(_both(A3, b3, solve.cg(precision="mixed", format=fmt).operator(A3)))


def test_noalloc_rejects_an_operator_beyond_the_scratch(dev):
    from jno_amd import _lib

    lib = _lib.load()
    small = torch.zeros(64, dtype=torch.float64, device="cuda")
    small32 = torch.zeros(64, dtype=torch.float32, device="cuda")
    op = _lib.JnoOp()
    op.format = _lib.JNO_FMT_DIA
    op.lanes = 1
    op.n = 1 << 22  # 4.2M rows: its mixed workspace is far above 64 MiB; nothing may be launched
    op.nnz = 0
    op.vals64 = small.data_ptr()
    op.vals32 = small32.data_ptr()
    op.ndiag = 1
    st = _lib.JnoPcgStats()
    need = lib.jno_pcg_workspace_size_op(C.byref(op), 1)
    assert need > (64 << 20)
    rc = lib.jno_pcg_mixed_noalloc(C.byref(op), _lib.ptr(small), _lib.ptr(small), _lib.ptr(small), 0, 1e-10, 0.0, 10,
                                   0.1, C.byref(st), _lib.stream_ptr())
    assert rc == _lib.JNO_ECAPACITY
    assert "static scratch" in _lib.last_error()
