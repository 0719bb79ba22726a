"""The benchmark configurations at their full sizes: 128^3 (precision), 64^3 (differentiation).

The FP64 reference here is the oracle's C + OpenMP restatement of the same Jacobi-PCG (jax cg
semantics, rtol 1e-10) run on the host, plus size-independent properties: closed-form nnz,
pattern SHA-256, symmetry, true FP64 residual <= 1e-10."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _sha(t):
    return hashlib.sha256(np.ascontiguousarray(_np(t), dtype="<i4").tobytes()).hexdigest()


def _kuhn_nnz(n):
    return (n + 1) ** 3 + 2 * (3 * n * (n + 1) ** 2 + 3 * n**2 * (n + 1) + n**3)


@pytest.mark.parametrize("n", [32, 64])
def test_pattern_sha256(dev, n):
    import jno_amd as J

    meta = json.load(open(os.path.join(GOLDEN, "pattern_sha256.json")))[f"p1_n{n}"]
    plan = J.kuhn_box(n).plan(1)
    assert plan.nnz == meta["nnz"] == _kuhn_nnz(n)
    assert _sha(plan.rowptr) == meta["rowptr_sha256"]
    assert _sha(plan.col) == meta["col_sha256"]


def _cpu_system(n, k):
    from oracle import cpu_fem

    p, t = O.kuhn_box(n)
    mask = np.any((p == 0.0) | (p == 1.0), axis=1).astype(np.uint8)
    return cpu_fem.P1System(p, t, k, 1.0, mask)


def test_c2_128_pattern_and_mixed_solve(dev):
    import jno_amd as J
    from jno_amd import precond, solve

    n = 128
    m = J.kuhn_box(n)
    plan = m.plan(1)
    N = m.n_nodes
    assert N == 2_146_689 and m.n_cells == 12_582_912
    assert plan.nnz == 31_802_497 == _kuhn_nnz(n)
    # rows sorted strictly, structurally symmetric
    rp = plan.rowptr.long()
    rows = torch.repeat_interleave(torch.arange(N, device=dev), rp[1:] - rp[:-1])
    col = plan.col.long()
    key = rows * N + col
    assert bool((key[1:] > key[:-1]).all())
    assert torch.equal(torch.sort(col * N + rows).values, key)
    assert int((m.boundary_mask() > 0).sum()) == 6 * n**2 + 2 == 98_306

    k_host = np.exp(0.5 * np.random.default_rng(0).standard_normal(m.n_cells))
    s = J.fem_poisson(m)
    A, b = s.evaluate({"k": torch.tensor(k_host, device=dev)})
    assert A.drop_zeros().nnz == 14_340_213
    mixed = solve.cg(tol=1e-10, precision="mixed")
    u = mixed(A, b, M=precond.jacobi())
    st = mixed.last_stats
    assert st["relres"] <= 1e-10, st
    # the FP64 GPU PCG on the same operator
    f64 = solve.cg(tol=1e-10, precision="fp64")
    u64 = f64(A, b, M=precond.jacobi())
    # jax cg semantics: the recursive residual meets rtol; the true one is reported beside it
    assert f64.last_stats["relres_recursive"] <= 1e-10 and f64.last_stats["relres"] <= 2e-10
    # FP64 CPU restatement of the reference solve (jax cg semantics, rtol 1e-10)
    cs = _cpu_system(n, k_host)
    assert np.abs(cs.vals - _np(A.vals)).max() <= 1e-12 * np.abs(cs.vals).max()
    x, it, rel, _ = cs.pcg(1e-10)
    assert abs(f64.last_stats["iterations"] - it) <= 2, (f64.last_stats, it)
    err_mixed, err_f64 = _rel(_np(u), x), _rel(_np(u64), x)
    print(f"\n128^3: mixed {st['iterations']} it ({st['refinements']} replacements) err {err_mixed:.2e}; "
          f"fp64 {f64.last_stats['iterations']} it err {err_f64:.2e}; cpu {it} it")
    assert err_f64 <= 1e-9 and err_mixed <= 1e-9


def test_c3_64_gradient(dev):
    import jno_amd as J
    from jno_amd import adjoint, precond, solve

    n = 64
    m = J.kuhn_box(n)
    T = m.n_cells
    assert T == 1_572_864
    s = J.fem_poisson(m)
    M = s.mass()
    k = np.exp(0.5 * np.random.default_rng(0).standard_normal(T))
    ks = np.exp(0.5 * np.random.default_rng(1).standard_normal(T))
    sv = solve.cg(tol=1e-10, precision="mixed")
    As, bs = s.evaluate({"k": torch.tensor(ks, device=dev)})
    ustar = sv(As, bs, M=precond.jacobi())
    kt = torch.tensor(k, device=dev, requires_grad=True)
    u = adjoint.solve(s, kt, sv)
    e = u - ustar
    L = 0.5 * torch.dot(e, M.mv(e))
    L.backward()
    g = _np(kt.grad)
    # oracle: FP64 PCG restatement at a tight tolerance + the closed-form contraction
    cs = _cpu_system(n, k)
    css = _cpu_system(n, ks)
    uo, _, _, _ = cs.pcg(1e-13)
    uso, _, _, _ = css.pcg(1e-13)
    p, t = O.kuhn_box(n)
    Mo = O.mass_matrix_p1(p, t)
    rhs = Mo @ (uo - uso)
    d = O.boundary_nodes(p, t)
    rhs[d] = 0.0
    cs.b = rhs
    lam, _, _, _ = cs.pcg(1e-13)
    go = O.adjoint_grad_k_p1(p, t, uo, lam, d)
    assert _rel(g, go) <= 1e-6, _rel(g, go)
    # central FD at h = 1e-6 k_e on 32 seeded elements.  The perturbed states are u0 + d with
    # A(k +- h) d = b - A(k +- h) u0 solved tightly (defect correction), and the loss difference is
    # formed from the increments, L+ - L- = e0^T M (d+ - d-) + (d+^T M d+ - d-^T M d-)/2, so the
    # 1e-6 relative change of one element out of 1.57M is not lost to cancellation in L itself.
    # A is linear in k, so A(k +- h) u0 = A(k) u0 +- h B_e u0 with B_e u0 formed once from
    # (A(k + k_e 1_e) - A(k)) u0 / k_e: the rounding of the assembled entries (eps |A_ij|, about
    # 1e-17, against the h K_e change of about 1e-9, and not free of the constant null space of
    # K_e) would otherwise put a noise of about 1e-6 / (h / k_e) of g_e into the difference.
    fine = solve.cg(tol=1e-12, precision="fp64")

    def tight(A_, b_):
        # one defect-correction pass: the error left by rtol 1e-12 (up to cond(A) * 1e-12) would
        # otherwise show as FD error on elements whose gradient is small
        x_ = fine(A_, b_, M=precond.jacobi())
        return x_ + fine(A_, b_ - A_.mv(x_), M=precond.jacobi())

    # the per-element comparison uses the same contraction (grad_k) on tightly solved states
    A0, b0 = s.evaluate({"k": kt.detach()})
    ustar = tight(As, bs)
    u0 = tight(A0, b0)
    e0 = u0 - ustar
    rhs = M.mv(e0)
    rhs[s.dmask.bool()] = 0.0
    g = _np(s.grad_k(u0, tight(A0, rhs)))
    r0 = b0 - A0.mv(u0)
    Au0 = A0.mv(u0)
    rng = np.random.default_rng(2024)
    errs = []
    for el in rng.choice(T, 32, replace=False):
        h = 1e-6 * k[el]
        kk = kt.detach().clone()
        kk[el] *= 2.0
        Bu = (s.evaluate({"k": kk})[0].mv(u0) - Au0) / k[el]
        ds = []
        for sg in (1, -1):
            kk = kt.detach().clone()
            kk[el] += sg * h
            A2, _ = s.evaluate({"k": kk})
            ds.append(tight(A2, r0 - (sg * h) * Bu))
        dp, dm = ds
        dL = torch.dot(e0, M.mv(dp - dm)) + 0.5 * (torch.dot(dp, M.mv(dp)) - torch.dot(dm, M.mv(dm)))
        fd = float(dL) / (2 * h)
        errs.append(abs(fd - g[el]) / abs(g[el]))
        assert abs(fd - g[el]) <= 1e-6 * abs(g[el]), (el, fd, g[el])
    print(f"\n64^3 gradient: rel L2 vs oracle {_rel(g, go):.2e}; FD rel err max {max(errs):.2e}, "
          f"median {float(np.median(errs)):.2e} over {len(errs)} elements")
