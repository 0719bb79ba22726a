"""DLPack interop and the BCOO -> CSR conversion of the linear= slot (host logic; no GPU needed)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from jno_amd.interop import as_tensor, csr_from_bcoo, to_csr


def test_as_tensor_takes_dlpack_exporters_without_copy():
    a = np.arange(12, dtype=np.float64)
    t = as_tensor(a)  # numpy arrays export __dlpack__
    assert isinstance(t, torch.Tensor) and t.dtype == torch.float64
    a[3] = -7.0
    assert t[3].item() == -7.0  # a view, not a copy
    u = torch.ones(3)
    assert as_tensor(u) is u


def test_csr_from_bcoo_sorts_and_sums_duplicates():
    rng = np.random.default_rng(0)
    n = 40
    rows = rng.integers(0, n, 300)
    cols = rng.integers(0, n, 300)
    # the reference's Dirichlet handling appends unit-diagonal triplets after the element ones
    d = np.array([0, 5, 17, 39])
    rows = np.concatenate([rows, d])
    cols = np.concatenate([cols, d])
    data = rng.standard_normal(rows.size)
    data[-4:] = 1.0
    idx = np.stack([rows, cols], axis=1).astype(np.int32)
    A = csr_from_bcoo(data, idx, (n, n))
    ref = sp.coo_matrix((data, (rows, cols)), shape=(n, n)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(A.rowptr.numpy(), ref.indptr)
    assert np.array_equal(A.col.numpy(), ref.indices)
    assert np.allclose(A.vals.numpy(), ref.data, rtol=1e-15, atol=1e-15)
    # a BCOO-carrying operator object gives the same CSR
    class Op:
        shape = (n, n)
        bcoo = (torch.tensor(data), torch.tensor(idx))

    B = to_csr(Op())
    assert torch.equal(B.rowptr, A.rowptr) and torch.equal(B.col, A.col) and torch.equal(B.vals, A.vals)
    # and so does a BCOO itself (.data, .indices, .shape, as JAX's BCOO and jno_amd.BCOO carry them)
    from jno_amd.fem import BCOO

    Cm = to_csr(BCOO(torch.tensor(data), torch.tensor(idx), (n, n)))
    assert torch.equal(Cm.rowptr, A.rowptr) and torch.equal(Cm.col, A.col) and torch.equal(Cm.vals, A.vals)


def test_csr_from_bcoo_rejects_bad_input():
    with pytest.raises(ValueError, match="square"):
        csr_from_bcoo(np.ones(2), np.array([[0, 0], [1, 1]]), (2, 3))
    with pytest.raises(ValueError, match="out of range"):
        csr_from_bcoo(np.ones(2), np.array([[0, 0], [2, 1]]), (2, 2))
    with pytest.raises(TypeError):
        to_csr(object())
