"""Benchmark of the jno_amd FEM hot path on one MI355X.

Workload (default): P1 Poisson -div(k grad u) = 1 on the unit cube, 128^3 hexes x 6 Kuhn tets
(12,582,912 tets, 2,146,689 nodes), k_e = exp(0.5 z_e) with z ~ N(0,1) from numpy
default_rng(0), zero Dirichlet on the boundary.  One step = re-assemble A(k), b (the parametric
``evaluate``) + solve to relative residual 1e-10 with FP32 Jacobi-PCG inner iterations and FP64
residual replacement, stopping on the true FP64 residual.  The CSR plan (pattern) is built once,
as the reference caches it.

Also reported: the plain FP64 GPU PCG on the same operator layout, the forward+adjoint gradient
rate on 64^3 (dL/dk for all 1,572,864 elements), the transient heat march, the C4 batch (256
uneven systems; 64 right-hand sides on 96^3) and the C5 P2 elasticity solve on the
element-by-element operator, each with its roofline fraction, the live HIP-event roofline of the
dominant kernel, and the CPU baseline (the C + OpenMP FP64 restatement, a bounded sample on the
host: one warm-up, then the median of 3).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-grad] [--no-batch] [--no-p2]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector and matrix)
REPLACE_FACTOR = 0.01
BATCH_REPLACE_FACTOR = 0.003  # the batch default (fewer replacements pay across 256 systems)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 steps (about 0.26 s at 128^3): a host stall of a few ms, seen now and then on these boxes,
    # moves a mean of 5 steps by a tenth of a millisecond
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="lattice cells per axis")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-transient", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-p2", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the bounded CPU PCG sample")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this path runs on one GPU (it would shard by batch instances or by mesh slabs "
              "with a halo exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import jno_amd as J
    from jno_amd import _lib, precond, solve
    from jno_amd.operator import SolverOperator

    dev = torch.device("cuda:0")
    n = args.n
    log(f"[bench] mesh {n}^3 + plan")
    mesh = J.kuhn_box(n, device=dev)
    T, N = mesh.n_cells, mesh.n_nodes
    k_host = np.exp(0.5 * np.random.default_rng(args.seed).standard_normal(T))
    k = torch.tensor(k_host, device=dev)
    system = J.fem_poisson(mesh)
    t0 = time.perf_counter()
    plan = mesh.plan(1)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0

    # FP64 residual replacement after a 100x drop of the FP32 residual: on this P1 Poisson operator
    # it takes 5 instead of 10 replacements at the same 645 inner iterations (the library default,
    # 0.1, is the safer choice for harder operators such as P2 elasticity)
    mixed = solve.cg(tol=1e-10, precision="mixed", replace_factor=REPLACE_FACTOR)
    fp64 = solve.cg(tol=1e-10, precision="fp64")

    def step(solver):
        A, b = system.evaluate({"k": k})
        u = solver(A, b, M=precond.jacobi())
        return A, b, u

    log("[bench] warmup")
    for _ in range(args.warmup):
        step(mixed)
    torch.cuda.synchronize()

    log(f"[bench] timed: {args.steps} steps (assemble + mixed solve to 1e-10)")
    lib = _lib.load()
    lib.jno_pcg_profile(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        A, b, u = step(mixed)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    prof_ms, prof_n = (C.c_double * 3)(), C.c_int64()
    lib.jno_pcg_profile_read(prof_ms, C.byref(prof_n))
    lib.jno_pcg_profile(0)
    st_mixed = dict(mixed.last_stats)
    op = mixed.operator(A)

    # component times (one extra, untimed-in-value step split into its parts).  Once the first solve
    # has picked DIA, the assembly writes the solver layout in the same pass (layout_fused); the
    # stand-alone layout from the CSR is timed beside it for reference (not part of the step then).
    torch.cuda.synchronize()
    ta = time.perf_counter()
    A2, b2 = system.evaluate({"k": k})
    torch.cuda.synchronize()
    t_assemble = time.perf_counter() - ta
    layout_fused = ("solver_op", "auto", True) in A2._cache
    ta = time.perf_counter()
    op2 = SolverOperator(A2, fmt=op.format, lanes=op.lanes, fp32=True)
    torch.cuda.synchronize()
    t_layout = time.perf_counter() - ta

    log("[bench] plain FP64 PCG on the same operator")
    fp64(A, b, M=precond.jacobi())  # warm
    torch.cuda.synchronize()
    ta = time.perf_counter()
    u64 = fp64(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - ta
    st_fp64 = dict(fp64.last_stats)
    agree = float(torch.linalg.norm(u - u64) / torch.linalg.norm(u64))

    # the plain FP64 PCG on the reference's own layout: CSR of the zero-dropped pattern (compress_eager)
    log("[bench] plain FP64 PCG on CSR")
    fp64csr = solve.cg(tol=1e-10, precision="fp64", format="csr")
    fp64csr(A, b, M=precond.jacobi())  # warm
    torch.cuda.synchronize()
    ta = time.perf_counter()
    fp64csr(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t_fp64csr = time.perf_counter() - ta
    st_fp64csr = dict(fp64csr.last_stats)

    # per-kernel live timing (HIP events on the solve's stream, every 8th iteration sampled) and the
    # algorithmic bytes each kernel must move per launch (FP32 inner iterations, 4-byte values)
    blk, cap, vec, mode = solve.get_tuning("mixed")
    nsamp = max(prof_n.value, 1)
    ms = [prof_ms[j] / nsamp for j in range(3)]
    if op.format == "dia":
        spmv_mat = N * 4 * op.ndiag  # main + stored upper diagonals (the lower ones are the same values)
    else:
        spmv_mat = op.nnz_stored * (4 + 4) + 4 * (N + 1)
    if mode == 5 and op.format == "dia":  # one kernel per iteration on the Jacobi-scaled operator
        # reads r, q, p, x and the stored upper diagonals; writes r, p, q, x.  k1_vector 1, 2 and 3
        # run the register-resident variants on the 4-diagonal operators with 2 or 3 near diagonals
        # (3: pipelined tiles, on an even grid)
        quad = vec in (1, 2, 3) and blk <= 512 and op.ndiag == 4 and 2 <= sum(o <= 256 for o in op.offsets[1:]) <= 3
        kname = ("pcg_one_p<dia, f32, scaled>" if vec == 3 else "pcg_one_q<dia, f32, scaled>") if quad \
            else "pcg_one<dia, f32, scaled>"
        if kname.startswith("pcg_one_p"):
            # x every other launch: a deferring launch moves r, q, p (24 B) and the diagonals, a
            # flushing one also reads x and p_{k-2} and writes x (12 B): 6 B per row on average
            kernels = [(kname, ms[1], N * 4 * (op.ndiag - 1) + N * (3 * 4 + 3 * 4) + N * 6)]
        else:
            kernels = [(kname, ms[1], N * 4 * (op.ndiag - 1) + N * (4 * 4 + 4 * 4))]
    elif mode == 4 and op.format == "dia":  # K1 K2 on the Jacobi-scaled operator: no dinv, unit diagonal
        k1 = "pcg_k1_lds<dia, f32, scaled>" if vec == 3 else "pcg_k1<dia, f32, scaled>"  # vec 3: p staged in LDS
        kernels = [(k1, ms[1], N * 4 * (op.ndiag - 1) + N * (3 * 4 + 3 * 4)),
                   ("pcg_k2<scaled>", ms[2], N * (2 * 4 + 4))]
    elif mode == 3 and op.format == "dia":  # K0 KS K2 on the Jacobi-scaled operator: no dinv, unit diagonal
        kernels = [("pcg_k0<scaled>", ms[0], N * (3 * 4 + 2 * 4)),
                   ("pcg_ks<dia, f32, unit>", ms[1], N * 4 * (op.ndiag - 1) + N * (4 + 4)),
                   ("pcg_k2<scaled>", ms[2], N * (2 * 4 + 4))]
    elif mode in (1, 3):  # K0 (p, x update), KS (q = A p, p.q), K2 (r update, r.z, r.r)
        kernels = [("pcg_k0", ms[0], N * (4 * 4 + 2 * 4)),
                   (f"pcg_ks<{op.format}, f32>", ms[1], spmv_mat + N * (4 + 4)),
                   ("pcg_k2", ms[2], N * (3 * 4 + 4))]
    elif mode == 2:  # KA (all vector updates), KB (w = A u, u.w)
        kernels = [("pcg_ka", ms[0], N * (6 * 4 + 5 * 4)),
                   (f"pcg_kb<{op.format}, f32>", ms[1], spmv_mat + N * (4 + 4))]
    else:  # modes 0 and 4 off DIA: K1 (p recompute for gathered neighbours, x update, q = A p), K2
        kernels = [(f"pcg_k1<{op.format}, f32>", ms[1], spmv_mat + N * (4 * 4 + 3 * 4)),
                   ("pcg_k2", ms[2], N * (3 * 4 + 4))]
    kernel_rows = {name: {"avg_ms": round(t, 5), "bytes": b, "gbs": round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None}
                   for name, t, b in kernels}
    dom_name, dom_ms, dom_bytes = max(kernels, key=lambda kv: kv[1])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = committed_traffic(dom_name, f"p1-{n}")
    iters = st_mixed["iterations"]
    out = {
        "metric": "P1 Poisson 128³ time-to-1e-10 + HBM GB/s; fwd+adjoint grads/s vs FP64 plain",
        "value": round(t_step, 6),
        "unit": "s per assemble+solve to rtol 1e-10 (128³ P1, mixed FP32/FP64)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (Kuhn lattice, lognormal k from numpy default_rng seed)",
        "config": {"workload": f"P1 Poisson {n}^3 Kuhn tets (T={T}, N={N}), k=exp(0.5 z) seed {args.seed}, "
                               "f=1, zero Dirichlet, rtol 1e-10", "parallelism": "single GPU"},
        "solve": {
            "mixed_s": round(st_mixed["seconds"], 6), "mixed_iterations": iters,
            "mixed_refinements": st_mixed["refinements"], "mixed_relres_true": st_mixed["relres"],
            "fp64_s": round(t_fp64, 6), "fp64_iterations": st_fp64["iterations"], "fp64_relres_true": st_fp64["relres"],
            "speedup_mixed_vs_fp64": round(t_fp64 / max(st_mixed["seconds"], 1e-12), 3),
            "fp64_csr_s": round(t_fp64csr, 6), "fp64_csr_iterations": st_fp64csr["iterations"],
            "fp64_csr_operator": st_fp64csr["operator"],
            "speedup_mixed_vs_fp64_csr": round(t_fp64csr / max(st_mixed["seconds"], 1e-12), 3),
            "mixed_vs_fp64_rel_l2": agree,
            "dof_iterations_per_s": N * iters / max(st_mixed["seconds"], 1e-12),
            "assemble_s": round(t_assemble, 6), "layout_fused": layout_fused,
            "layout_from_csr_s": round(t_layout, 6), "plan_s": round(t_plan, 6),
            "operator": op.describe() | {"nnz_struct": plan.nnz},
        },
        "roofline": {
            "bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms, "sampled_iterations": prof_n.value,
            "launch_shape": {"block": blk, "grid_cap": cap, "spmv_variant": vec, "mode": mode},
            "kernels": kernel_rows,
        },
    }

    if not args.no_grad:
        log("[bench] forward + adjoint gradient on 64^3")
        out["gradient"] = grad_rate(J, torch, dev)

    if not args.no_transient:
        log(f"[bench] transient heat march on {n}^3")
        out["transient"] = heat_rate(J, torch, dev, n)

    if not args.no_batch:
        log("[bench] C4 batch: 256 uneven systems, 64 right-hand sides on 96^3")
        out["batch"] = c4_rate(J, torch, dev)

    if not args.no_p2:
        log("[bench] C5 P2 elasticity 32^3 on the element-by-element operator")
        out["p2_elasticity"] = c5_rate(J, torch, dev)

    if not args.no_cpu:
        log("[bench] CPU baseline (C + OpenMP FP64 restatement, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(n, args.seed, args.cpu_seconds, st_fp64["iterations"])
    print(json.dumps(out), flush=True)


def committed_traffic(kernel, workload):
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the same counter on a copy kernel of known size), or
    None when no summary covers them."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if not os.path.isdir(path):
        return None
    for fn in sorted(os.listdir(path)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(path, fn)):
            parts = line.strip().strip("`").split()
            if len(parts) == 4 and parts[0] == "TRAFFIC" and parts[1] == kernel.replace(" ", "") and parts[2] == workload:
                return int(parts[3])
    return None


def grad_rate(J, torch, dev, n=64, reps=5):
    from jno_amd import adjoint, solve

    mesh = J.kuhn_box(n, device=dev)
    T = mesh.n_cells
    system = J.fem_poisson(mesh)
    kt = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(T)), device=dev)
    ks = torch.tensor(np.exp(0.5 * np.random.default_rng(1).standard_normal(T)), device=dev)
    M = system.mass()
    res = {}
    for prec in ("mixed", "fp64"):
        sv = solve.cg(tol=1e-10, precision=prec, replace_factor=REPLACE_FACTOR)
        from jno_amd import precond

        Ast, bst = system.evaluate({"k": ks})
        ustar = sv(Ast, bst, M=precond.jacobi())

        def once():
            k = kt.clone().requires_grad_(True)
            u = adjoint.solve(system, k, sv)
            e = u - ustar
            L = 0.5 * torch.dot(e, M.mv(e))
            L.backward()
            return k.grad

        once()
        # each repetition timed on its own: the rate is the median's (a host stall of tens of ms,
        # seen now and then on these boxes, would dominate a mean of a few 8 ms gradients)
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g = once()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        res[f"{prec}_grads_per_s"] = round(1.0 / dt, 3)
        res[f"{prec}_s_per_grad"] = round(dt, 6)
        res[f"{prec}_s_per_grad_mean"] = round(float(np.mean(times)), 6)
        res[f"{prec}_grad_norm"] = float(torch.linalg.norm(g))
    res["workload"] = (f"P1 Poisson {n}^3 (T={T}), L = 1/2 (u-u*)^T M (u-u*), dL/dk for all elements; "
                       f"median of {reps} timed gradients after one warm-up")
    res["speedup_mixed_vs_fp64"] = round(res["fp64_s_per_grad"] / res["mixed_s_per_grad"], 3)
    return res


def heat_rate(J, torch, dev, n=128, steps=20):
    """Backward-Euler steps of the transient heat block u_t v + k grad u . grad v on the unit cube
    (the 3-D form of the reference's heat benchmark): lognormal k (seed 0), u0 = sin sin sin, u = 0
    on the boundary, dt = 0.2/1500, each step solved to rtol 1e-10 with warm start."""
    from jno_amd import solve
    from jno_amd.transient import default_step_solver

    mesh = J.kuhn_box(n, device=dev)
    system = J.fem_poisson(mesh, f=0.0)
    k = torch.tensor(np.exp(0.5 * np.random.default_rng(0).standard_normal(mesh.n_cells)), device=dev)
    blk = J.SemidiscreteTimeBlock.from_system(system, {"k": k}, dt=0.2 / 1500)
    p = mesh.points
    u0 = (torch.sin(np.pi * p[:, 0]) * torch.sin(np.pi * p[:, 1]) * torch.sin(np.pi * p[:, 2])).contiguous()
    res = {"workload": f"heat {n}^3 P1, theta = 1, dt = 0.2/1500, {steps} steps to rtol 1e-10 (warm start)"}
    finals = {}
    for name, sv in (("mixed", default_step_solver()), ("fp64", solve.cg(tol=1e-10))):
        blk.march(u0, 2, linear_solve=sv)  # warm: step operators, layouts, code objects
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        finals[name] = blk.march(u0, steps, linear_solve=sv)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        its = blk.last_iterations
        res[f"{name}_ms_per_step"] = round(dt * 1e3, 4)
        res[f"{name}_iterations_per_step"] = round(sum(its) / len(its), 2)
        res["step_operator"] = sv.last_stats["operator"]
    res["mixed_vs_fp64_rel_l2"] = float((finals["mixed"] - finals["fp64"]).norm() / finals["fp64"].norm())
    res["speedup_mixed_vs_fp64"] = round(res["fp64_ms_per_step"] / res["mixed_ms_per_step"], 3)
    # a time-dependent coefficient (1 + 50 t) k: the operator re-assembled at every step's end time,
    # the step and explicit operators re-formed on their diagonal layouts
    blk_t = J.SemidiscreteTimeBlock.from_system(system, {"k": k}, dt=0.2 / 1500,
                                                coefficient_fn=lambda t, a: (1.0 + 50.0 * t) * k)
    sv = default_step_solver()
    blk_t.march(u0, 3, linear_solve=sv)  # warm: the DIA offsets of the plan, the mass's layout
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    blk_t.march(u0, steps, linear_solve=sv)
    torch.cuda.synchronize()
    its = blk_t.last_iterations
    res["time_dependent_mixed_ms_per_step"] = round((time.perf_counter() - t0) / steps * 1e3, 4)
    res["time_dependent_iterations_per_step"] = round(sum(its) / len(its), 2)
    return res


def cpu_baseline(n, seed, budget_s, full_iters):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import cpu_fem
    from oracle import fem_oracle as O

    p, t = O.kuhn_box(n)
    k = np.exp(0.5 * np.random.default_rng(seed).standard_normal(t.shape[0]))
    mask = np.any((p == 0.0) | (p == 1.0), axis=1).astype(np.uint8)
    s = cpu_fem.P1System(p, t, k, 1.0, mask)
    # warm-up (the construction's assembly and a short probe, which also sizes the bounded sample to
    # the budget), then three timed repetitions of assembly + the PCG sample: medians
    _, it0, _, dt0 = s.pcg(1e-10, maxiter=10)
    per_it = dt0 / max(it0, 1)
    sample = int(max(20, min(full_iters, budget_s / 3 / max(per_it, 1e-9))))
    t_asm, t_it = [], []
    for _ in range(3):
        t_asm.append(s.assemble())
        _, its, _, dts = s.pcg(1e-10, maxiter=sample)
        t_it.append(dts / max(its, 1))
    asm, per_it = float(np.median(t_asm)), float(np.median(t_it))
    est = asm + per_it * full_iters
    return {
        "value": round(est, 3),
        "unit": "s per assemble+solve to rtol 1e-10 (estimated: assembly + per-iteration time x full iteration count)",
        "cores": cpu_fem.threads(),
        "kind": "port",
        "sample": f"C+OpenMP FP64 restatement on the structural CSR (nnz {len(s.col)}), one warm-up then the median "
                  f"of 3: assembly {asm:.3f} s (pattern {s.t_pattern:.3f} s untimed) + {its} of {full_iters} PCG "
                  f"iterations at {per_it * 1e3:.2f} ms/iteration",
    }


def c5_rate(J, torch, dev, n=32):
    """C5: P2 elasticity on the n^3 Kuhn box (E = exp(0.3 z), default_rng(0); nu = 0.3; clamp x = 0,
    traction (0, 0, -1) on x = 1), mixed solve to rtol 1e-10 on the element-by-element operator, with
    the assembled BSR3 layout beside it; the EBE SpMV (element pass + node gather) timed live."""
    from jno_amd import _lib, precond, solve
    from jno_amd.elasticity import ElasticitySystem

    lib = _lib.load()
    s = ElasticitySystem(J.kuhn_box(n, device=dev))
    T, N = s.mesh.n_cells, s.mesh.n_nodes
    E = torch.tensor(np.exp(0.3 * np.random.default_rng(0).standard_normal(T)), device=dev)
    A, b = s.element_operator({"E": E})
    sv = solve.cg(tol=1e-10, precision="mixed")
    sv(A, b, M=precond.jacobi())  # warm: code objects, diagonal
    torch.cuda.synchronize()
    lib.jno_pcg_profile(1)
    t0 = time.perf_counter()
    u = sv(A, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t_ebe = time.perf_counter() - t0
    prof_ms, prof_n = (C.c_double * 3)(), C.c_int64()
    lib.jno_pcg_profile_read(prof_ms, C.byref(prof_n))
    lib.jno_pcg_profile(0)
    st = dict(sv.last_stats)
    g = A.geometry
    slots, rows = g.n_slots, 3 * N
    if A.structured:
        # the half-lattice SpMV (one kernel): x once, E per cell, the point -> node map and the
        # constrained-dof bits once, y written; 900 fused multiply-adds per cell
        spmv_kernel = "ebe_box_k<f32> (one SpMV on the half lattice)"
        spmv_bytes = rows * 4 + T * 4 + N * (4 + 1) + rows * 4
    else:
        # element pass (slot cells, bits, scale, incidence positions, x once, products written) +
        # gather (node pointers, products, x, mask, y)
        spmv_kernel = "ebe_elem_k<f32> + ebe_gather_k<f32> (one SpMV)"
        spmv_bytes = slots * (40 + 4 + 4 + 40) + rows * 4 + T * 30 * 4 + (N + 1) * 4 + T * 30 * 4 + rows * (4 + 1 + 4)
    spmv_flops = T * 900 * 2
    spmv_ms = prof_ms[1] / max(prof_n.value, 1)
    # the general element-by-element kernels on the same operator, for comparison
    Ag, _ = s.element_operator({"E": E}, structured=False)
    svg = solve.cg(tol=1e-10, precision="mixed")
    svg(Ag, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    svg(Ag, b, M=precond.jacobi())
    torch.cuda.synchronize()
    t_general = time.perf_counter() - t0
    del Ag
    # the forward solve plus the gradient of the compliance C = b.u with respect to every element's
    # E (autograd through the closed-form contraction), on the element-by-element operator
    from jno_amd.elasticity import compliance

    Eg = E.clone().requires_grad_(True)
    Cv, _ = compliance(s, Eg, sv)
    Cv.backward()  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Eg = E.clone().requires_grad_(True)
    Cv, _ = compliance(s, Eg, sv)
    Cv.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    g_E = Eg.grad
    # C falls as any E grows (K_e is positive semidefinite), up to rounding of a near-zero quadratic form
    grad_finite = bool(torch.isfinite(g_E).all()) and bool((g_E <= 1e-12 * float(g_E.abs().max())).all())
    Ab, bb = s.evaluate({"E": E})
    # the assembled operator (MFMA element matrices, colored scatter through the pair table)
    ta = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.evaluate({"E": E})
        torch.cuda.synchronize()
        ta.append(time.perf_counter() - t0)
    t_assemble = float(np.median(ta))
    # the assembled operator on the layout and lanes the autotuner picks (BSR3 at this size)
    svb = solve.cg(tol=1e-10, precision="mixed")
    svb(Ab, bb, M=precond.jacobi())  # warm: layout
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ub = svb(Ab, bb, M=precond.jacobi())
    torch.cuda.synchronize()
    t_bsr3 = time.perf_counter() - t0
    gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else None
    return {
        "workload": f"P2 elasticity {n}^3 (T={T}, {rows} dofs), E=exp(0.3 z) seed 0, nu=0.3, mixed to rtol 1e-10",
        "ebe_mixed_s": round(t_ebe, 6), "ebe_iterations": st["iterations"], "ebe_relres_true": st["relres"],
        "ebe_structured": bool(A.structured), "ebe_general_kernels_mixed_s": round(t_general, 6),
        "assembly_s": round(t_assemble, 6),
        "bsr3_mixed_s": round(t_bsr3, 6), "bsr3_iterations": svb.last_stats["iterations"],
        "assembled_operator": svb.last_stats.get("operator"),
        "speedup_ebe_vs_bsr3": round(t_bsr3 / t_ebe, 3),
        "ebe_vs_bsr3_rel_l2": float(torch.linalg.norm(u - ub) / torch.linalg.norm(ub)),
        "compliance_grad_s": round(t_grad, 6), "compliance_grad_nonpositive": grad_finite,
        "roofline": {"kernel": spmv_kernel, "avg_launch_ms": round(spmv_ms, 5),
                     "bytes_per_launch": spmv_bytes, "achieved": round(gbs, 1) if gbs else None, "unit": "GB/s",
                     "peak": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None,
                     "flops_per_launch": spmv_flops,
                     "fp32_tflops": round(spmv_flops / (spmv_ms * 1e-3) / 1e12, 2) if spmv_ms > 0 else None,
                     "fp32_peak_tflops": FP32_PEAK_TFLOPS,
                     "fp32_frac": round(spmv_flops / (spmv_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4) if spmv_ms > 0 else None,
                     # the peak without packed FP32 instructions, which this build does not emit
                     "fp32_unpacked_peak_tflops": FP32_PEAK_TFLOPS / 2,
                     "fp32_unpacked_frac": round(spmv_flops / (spmv_ms * 1e-3) / 1e12 / (FP32_PEAK_TFLOPS / 2), 4)
                     if spmv_ms > 0 else None,
                     "sampled_iterations": prof_n.value},
    }


def c4_rate(J, torch, dev, count=256, seed=4, n_rhs=64, n_rhs_mesh=96):
    """C4: ``count`` P1 Poisson systems on n^3 Kuhn boxes, n uniform in [24, 64] (default_rng(seed)),
    each with its own lognormal k (default_rng(1000 + i)), f = 1, solved together to rtol 1e-10; and
    ``n_rhs`` N(0,1) right-hand sides on one n_rhs_mesh^3 operator.  Effective bandwidth: the
    algorithmic bytes of every launch an item took part in, against the solve time: per row 42 B
    per one-kernel iteration on average (r, q, p read and written: 24; the 3 stored upper diagonals
    of the Jacobi-scaled operator: 12; xs read and written and p_{k-2} read every other launch: 6),
    and per residual replacement 32 B for folding xs into x64 (xs, x64
    read and written, dinv read) plus 60 B for the FP64 true residual (b, x64, 4 FP64 diagonals,
    dinv read; r written).  Each item runs iterations + replacements + 1 iteration launches (the
    resumed launch after each replacement and after the start)."""
    from jno_amd.batch import BatchSolver

    ns = np.random.default_rng(seed).integers(24, 65, size=count)
    systems = []
    for i, nn in enumerate(ns):
        m = J.kuhn_box(int(nn), device=dev)
        k = np.exp(0.5 * np.random.default_rng(1000 + i).standard_normal(m.n_cells))
        systems.append(J.fem_poisson(m).evaluate({"k": torch.tensor(k, device=dev)}))
    # the launch shape (rows per entry, workgroups, steps per graph replay) picked by the autotuner on
    # the first call, which also warms the layouts, the graph capture and the code objects
    bs = BatchSolver(tol=1e-10, autotune=True, replace_factor=BATCH_REPLACE_FACTOR)
    bs(systems)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs = bs(systems)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # the error against FP64 single solves of three of the systems (the largest, the smallest, the first)
    from jno_amd import precond, solve

    fp64 = solve.cg(tol=1e-10, precision="fp64")
    errs = []
    for i in sorted({0, int(np.argmax(ns)), int(np.argmin(ns))}):
        A, b = systems[i]
        ref = fp64(A, b, M=precond.jacobi())
        errs.append(float(torch.linalg.norm(xs[i] - ref) / torch.linalg.norm(ref)))
    its = [s["iterations"] for s in bs.last_stats]
    refs = [s["refinements"] for s in bs.last_stats]
    rows = [A.shape[0] for A, _ in systems]

    def batch_bytes(r, it, rf):
        return r * (42 * (it + rf + 1) + (32 + 60) * (rf + 1))

    moved = sum(batch_bytes(r, it, rf) for r, it, rf in zip(rows, its, refs))
    m96 = J.kuhn_box(n_rhs_mesh, device=dev)
    k96 = np.exp(0.5 * np.random.default_rng(7).standard_normal(m96.n_cells))
    A96, _ = J.fem_poisson(m96).evaluate({"k": torch.tensor(k96, device=dev)})
    rng = np.random.default_rng(8)
    rhs = [(A96, torch.tensor(rng.standard_normal(A96.shape[0]), device=dev)) for _ in range(n_rhs)]
    br = BatchSolver(tol=1e-10, autotune=True, replace_factor=BATCH_REPLACE_FACTOR)
    br(rhs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br(rhs)
    torch.cuda.synchronize()
    t_rhs = time.perf_counter() - t0
    its_r = [s["iterations"] for s in br.last_stats]
    moved_r = sum(batch_bytes(A96.shape[0], it, s["refinements"]) for it, s in zip(its_r, br.last_stats))
    gbs = moved / t_batch / 1e9
    return {
        "workload": f"{count} P1 Poisson systems, n^3 Kuhn boxes with n in [24, 64] (seed {seed}), {sum(rows)} rows; "
                    f"plus {n_rhs} N(0,1) right-hand sides on {n_rhs_mesh}^3; mixed to rtol 1e-10",
        "batch_s": round(t_batch, 6), "iterations_min_max": [min(its), max(its)],
        "launch_shape": bs.last_launch_shape, "rhs_launch_shape": br.last_launch_shape,
        "rel_l2_vs_fp64_single_solves": errs,
        "worst_relres": max(s["relres"] for s in bs.last_stats),
        "rhs_s": round(t_rhs, 6), "rhs_iterations_min_max": [min(its_r), max(its_r)],
        "one_kernel": bs.last_one_kernel,
        "roofline": {"kernel": "batch_one_k<256> (one kernel per step, quads), whole solve",
                     "bytes": moved, "achieved": round(gbs, 1), "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "rhs_achieved": round(moved_r / t_rhs / 1e9, 1)},
    }


if __name__ == "__main__":
    main()
