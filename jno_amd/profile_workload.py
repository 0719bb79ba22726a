"""A bounded, repeatable workload for profiler runs (rocprofv3 kernel traces and PMC counters).

    python -m jno_amd.profile_workload --case c2 --iters 40 [--precision mixed|fp64] [--copy-mib 1024]
        [--tuning block,grid_cap,k1_vector,mode] [--format auto|dia|bsr3|csr|ebe]

Builds the configuration's system (128^3 P1 Poisson for c2, 64^3 for c3, 32^3 P2 elasticity for
c5; --format ebe: its element-by-element operator), then runs exactly --iters Jacobi-PCG iterations (no convergence test, no residual guard), and
finally copies a buffer of --copy-mib MiB with a device-to-device copy kernel: the copy's
FETCH_SIZE/WRITE_SIZE give the counter-to-bytes calibration for the PCG kernels' counters.
--case c4 instead solves 256 uneven P1 systems (n in [24, 64]) to 1e-10 with the batched solver
(--batch-step one|three: the one-kernel step, the default, or the three-kernel step);
--case heat marches --iters backward-Euler steps of the n^3 transient heat block (--n, default 128).
--case c2step runs --iters whole C2 steps (128^3 assemble + mixed solve to 1e-10 with residual
replacement after a 100x drop of the FP32 residual), after two warm steps.
--case assemble re-assembles the n^3 P1 system (--n, default 128) --iters times on the structured-box
kernel and then --iters times on the incidence-gather kernel, both writing the DIA layout in the
same pass (one solve first picks the offsets).
"""
from __future__ import annotations

import argparse
import time

import numpy as np
import torch


def build(case: str, dev, fmt: str = "auto"):
    import jno_amd as J

    if case in ("c2", "c3"):
        n = 128 if case == "c2" else 64
        m = J.kuhn_box(n, device=dev)
        s = J.fem_poisson(m)
        k = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(m.n_cells)), device=dev)
        return s.evaluate({"k": k})
    if case == "c5":
        from jno_amd.elasticity import ElasticitySystem

        s = ElasticitySystem(J.kuhn_box(32, device=dev))
        E = torch.tensor(np.exp(0.3 * np.random.default_rng(0).standard_normal(s.mesh.n_cells)), device=dev)
        if fmt == "ebe":  # the element-by-element operator instead of an assembled layout
            return s.element_operator({"E": E})
        return s.evaluate({"E": E})
    raise ValueError(f"unknown case {case!r}")


def build_batch(dev, count=256, seed=4):
    """``count`` P1 Poisson systems on n^3 Kuhn boxes, n uniform in [24, 64] from ``seed``, each
    with its own lognormal k (default_rng(1000 + i))."""
    import jno_amd as J

    ns = np.random.default_rng(seed).integers(24, 65, size=count)
    systems = []
    for i, n in enumerate(ns):
        m = J.kuhn_box(int(n), device=dev)
        k = np.exp(0.5 * np.random.default_rng(1000 + i).standard_normal(m.n_cells))
        systems.append(J.fem_poisson(m).evaluate({"k": torch.tensor(k, device=dev)}))
    return systems


def run_heat(dev, steps: int, precision: str, n: int = 128, replace_factor=None):
    """3-D analogue of the reference's transient heat benchmark (benchmarks/fem_bench.py:241-259:
    u_t v + k grad u . grad v on the unit square, IC sin sin, u = 0 on the boundary, T = 0.2 in 1500
    backward-Euler steps): the unit cube at n^3 cells, lognormal k, IC sin sin sin, dt = 0.2/1500,
    ``steps`` steps each solved to rtol 1e-10 with warm start."""
    import jno_amd as J

    m = J.kuhn_box(n, device=dev)
    s = J.fem_poisson(m, f=0.0)
    k = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(m.n_cells)), device=dev)
    blk = J.SemidiscreteTimeBlock.from_system(s, {"k": k}, dt=0.2 / 1500)
    p = m.points
    u0 = (torch.sin(np.pi * p[:, 0]) * torch.sin(np.pi * p[:, 1]) * torch.sin(np.pi * p[:, 2])).contiguous()
    from jno_amd.transient import default_step_solver

    if precision == "mixed" and replace_factor is None:
        sv = default_step_solver()
    else:
        sv = J.solve.cg(tol=1e-10, precision=precision, replace_factor=replace_factor or 0.1)
    replace_factor = sv.replace_factor
    blk.march(u0, 2, linear_solve=sv)  # warm (step operators, layout, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = blk.march(u0, steps, linear_solve=sv)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its = blk.last_iterations
    print(f"heat {n}^3 {precision} rf {replace_factor:g}: {steps} steps in {dt * 1e3:.1f} ms ({dt / steps * 1e3:.3f} ms/step), "
          f"iterations/step {min(its)}..{max(its)} (mean {sum(its) / len(its):.1f}), |u_N| {float(u.norm()):.6e}, "
          f"operator {sv.last_stats['operator']}")


def run_assemble(dev, iters: int, n: int = 128):
    import jno_amd as J
    from jno_amd import precond

    m = J.kuhn_box(n, device=dev)
    k = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(m.n_cells)), device=dev)
    for structured in (True, False):
        s = J.fem_poisson(m, structured=structured)
        A, b = s.evaluate({"k": k})
        J.solve.cg(tol=1e-6)(A, b, M=precond.jacobi())  # picks the DIA offsets of the plan
        A, b = s.evaluate({"k": k})  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            A, b = s.evaluate({"k": k})
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / iters
        fused = A._cache.get(("solver_op", "auto", True))
        print(f"assemble {n}^3 {'structured box' if structured else 'incidence gather'}: {dt * 1e6:.1f} us per "
              f"evaluate, nnz {A.nnz}, DIA fused {fused is not None and fused.format == 'dia'}")


def run_c2_steps(dev, steps: int):
    import jno_amd as J
    from jno_amd import precond, solve

    m = J.kuhn_box(128, device=dev)
    system = J.fem_poisson(m)
    k = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(m.n_cells)), device=dev)
    sv = solve.cg(tol=1e-10, precision="mixed", replace_factor=0.01)
    for _ in range(2):  # warm: the layout choice, code objects
        A, b = system.evaluate({"k": k})
        sv(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        A, b = system.evaluate({"k": k})
        sv(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(steps, 1)
    st = sv.last_stats
    print(f"c2 step: {dt * 1e3:.3f} ms per assemble + mixed solve, {st['iterations']} inner launches, "
          f"{st.get('refinements')} replacements")


def calibration_copy(dev, mib: int):
    """A device-to-device copy of ``mib`` MiB: its FETCH_SIZE / WRITE_SIZE per byte calibrate the
    counters of the other kernels."""
    nbytes = mib << 20
    src = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"calibration copy: {nbytes} bytes read + {nbytes} bytes written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="c2")
    ap.add_argument("--iters", type=int, default=40)
    ap.add_argument("--precision", default="mixed", choices=("mixed", "fp64"))
    ap.add_argument("--format", default="auto")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--n", type=int, default=128, help="cells per axis of the heat case")
    ap.add_argument("--no-pace", action="store_true", help="fixed chunks for the one-kernel inner loops")
    ap.add_argument("--replace-factor", type=float, default=None,
                    help="residual replacement factor of the heat case (default: the step solver's)")
    ap.add_argument("--batch-step", default="one", choices=("one", "three"),
                    help="c4: one kernel per step (default) or the three-kernel step")
    ap.add_argument("--tuning", default=None,
                    help="launch shape block,grid_cap,k1_vector,mode of the iterations (default: the library's)")
    args = ap.parse_args()
    from jno_amd import precond
    from jno_amd.solve import LinearSolver

    dev = torch.device("cuda:0")
    if args.case == "c4":
        from jno_amd.batch import BatchSolver

        systems = build_batch(dev)
        bs = BatchSolver(tol=1e-10, steps_per_graph=8, one_kernel=args.batch_step == "one")
        bs(systems)  # warm (graphs, code objects)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bs(systems)
        torch.cuda.synchronize()
        its = [st["iterations"] for st in bs.last_stats]
        refs = [st["refinements"] for st in bs.last_stats]
        print(f"c4 batch ({'one kernel' if bs.last_one_kernel else 'three kernels'} per step): {len(systems)} "
              f"systems, {sum(A.shape[0] for A, _ in systems)} rows, {(time.perf_counter() - t0) * 1e3:.1f} ms, "
              f"iterations {min(its)}..{max(its)} (sum {sum(its)}), refinements {min(refs)}..{max(refs)}")
        calibration_copy(dev, args.copy_mib)
        return
    if args.no_pace:
        from jno_amd.solve import set_pacing

        set_pacing(False)
    if args.tuning:
        from jno_amd.solve import set_tuning

        blk, cap, vec, mode = (int(v) for v in args.tuning.split(","))
        set_tuning(args.precision, blk, cap, vec, mode)
    if args.case == "heat":
        run_heat(dev, args.iters, args.precision, args.n, args.replace_factor)
        return
    if args.case == "c2step":
        run_c2_steps(dev, args.iters)
        calibration_copy(dev, args.copy_mib)
        return
    if args.case == "assemble":
        run_assemble(dev, args.iters, args.n)
        calibration_copy(dev, args.copy_mib)
        return
    A, b = build(args.case, dev, args.format)
    # replace_factor tiny: no residual replacement inside the bounded run
    sv = LinearSolver("cg", 1e-30, 0.0, args.iters, precision=args.precision, format=args.format,
                      replace_factor=1e-30, guard=False)
    sv(A, b, M=precond.jacobi())  # warm (layout, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sv(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"{args.case} {args.precision}: {sv.last_stats['iterations']} iterations in {dt * 1e3:.3f} ms "
          f"({dt / max(sv.last_stats['iterations'], 1) * 1e6:.1f} us/iteration), operator {sv.last_stats['operator']}")
    calibration_copy(dev, args.copy_mib)


if __name__ == "__main__":
    main()
