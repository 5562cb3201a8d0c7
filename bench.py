#!/usr/bin/env python3
"""Benchmark: residual + Jacobian assembly of the nonlinear Q2 thermal problem.

Workload (one GPU): [0,1]^3, 96^3 Q2 hexahedra (193^3 = 7,189,057 DOFs,
454,756,609 CRS nonzeros), 3x3x3 Gauss points, k(T,x) = k0_e (1 + T^2) with
k0_e = exp(0.5 z_e) (z_e ~ N(0,1), seed 2024), source f_e ~ U[0,1] (seed 2025),
T = 0 on x = 0 and x = 1.  A step is one full J+R assembly (eight colour
launches of the fused kernel; the first element to touch an entry stores it,
Dirichlet rows included) with u, res and J resident in HBM.

Prints one JSON line: assembly MDOF/s plus the roofline of the dominant kernel
(algorithmic bytes B_asm per launch / its mean HIP-event duration) and a CPU
baseline (the oracle, single thread, on a 24^3 sample scaled per element);
beside them the GMRES iteration rate on the assembled Jacobian and the
batched-subgrid solve rate.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12  # 256 CUs x 128 B/clk x 2.4 GHz

# HBM traffic of k_thermal<jacres> (96^3 Q2, element colouring), per launch,
# from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this bench (mean of the
# 8 colour launches), scaled by the factors that make the same counters read
# the known size of the plain copy kernel: FETCH x 2048 B, WRITE x 1024 B.
# Source: profiles/k_thermal_q2hex_96cubed.md.  Only valid for that workload.
PMC_TRAFFIC = {(96, 0): {"fetch_kib": 529038.2, "write_kib": 1024350.0, "fetch_scale": 2048, "write_scale": 1024}}


def c2_problem(n: int):
    from mrhyde_amd.abi import MHX_RNG_LOGNORMAL, MHX_RNG_UNIFORM, Problem, rng_fill

    ne = n ** 3
    k0 = rng_fill(2024, ne, MHX_RNG_LOGNORMAL, 0.0, 0.5)
    f = rng_fill(2025, ne, MHX_RNG_UNIFORM, 0.0, 1.0)
    return Problem(dim=3, order=2, quadrature=4, n=(n, n, n), k0=k0, src_elem=f, diff_u2=1.0,
                   dbc_sides=1 | 2)


def b_asm(ne: int, n_e: int, n_c: int, n_vert: int, dim: int) -> float:
    """Algorithmic bytes of one coloured J+R assembly (element-matrix read+write,
    u gather, residual read+write, coefficients, vertex coordinates)."""
    return ne * (16 * n_e * n_e + 24 * n_e + 8 * n_c) + 8 * dim * n_vert


def cpu_baseline(n_full: int, sample_n: int = 24):
    from mrhyde_amd.abi import rng_fill
    from oracle import Oracle

    p = c2_problem(sample_n)
    o = Oracle(p)
    u = rng_fill(7, o.n_dof, 0, 0.0, 1.0)
    o.setDirichlet(u)
    t0 = time.perf_counter()
    o.assembleJacRes(u)
    dt = time.perf_counter() - t0
    ne_s, ne_f = sample_n ** 3, n_full ** 3
    n_full_dof = (2 * n_full + 1) ** 3
    t_full = dt * ne_f / ne_s
    return {"value": n_full_dof / t_full / 1e6, "unit": "MDOF/s", "cores": 1, "kind": "port",
            "sample": f"oracle J+R assembly of the same problem on a {sample_n}^3 Q2 mesh ({ne_s} elements, "
                      f"{dt:.2f} s), scaled per element to {n_full}^3"}


def krylov_bench(g, iters: int):
    """GMRES(50) with Jacobi on the assembled Jacobian (device-resident), a
    fixed number of iterations: all-FP64, and FP32 Jacobian / vectors / inner
    GMRES inside FP64 refinement.  Per-iteration HIP-event time against the
    algorithmic bytes of an average GMRES(m) iteration,
    B_spmv + w N (m + 3) with w the vector width."""
    from mrhyde_amd.abi import MHX_BUF_DU

    N, nnz, m = g.n_dof, g.nnz, 50
    out = {"workload": f"GMRES({m}) + Jacobi on the assembled Jacobian, {iters} iterations from du = 0"}
    for name, prec, wv, wa in (("fp64", "jacobi", 8, 8), ("fp32-inner", "jacobi-fp32", 4, 4)):
        g.state_upload(MHX_BUF_DU, np.zeros(N))
        g.synchronize()
        g.profile(True)
        g.profile_reset()
        t0 = time.perf_counter()
        st = g.solve_resident(solver="gmres", restart=m, tol=1e-14, max_iters=iters, precond=prec)
        g.synchronize()
        dt = time.perf_counter() - t0
        prof = g.profile_read()
        g.profile(False)
        it_ms, it_n = prof.get("Krylov::gmres_iteration", (float("nan"), 0))
        per_it = it_ms / max(it_n, 1)
        b_it = (wa + 4) * nnz + 4 * (N + 1) + 2 * wv * N + wv * N * (m + 3)
        out[name] = {"iterations": st["iters"], "refinements": st["refinements"], "rel_res": st["rel_res"],
                     "wall_s": round(dt, 3), "ms_per_iteration": round(per_it, 4),
                     "bytes_per_iteration": b_it, "achieved_GBps": round(b_it / (per_it * 1e-3) / 1e9, 1)}
    return out


def subgrid_bench(steps: int, per_call_sample: int = 16):
    """4096 independent Q1 thermal subproblems (n in {4,6,8,12,16}, lognormal
    k with sigma 1, Dirichlet data U[-1,1]) assembled and solved to 1e-10 in
    one persistent launch replayed from a hipGraph; FP64 and FP32-inner /
    FP64-refined; plus the plain per-call engine path timed on a sample."""
    from mrhyde_amd import Mhx, SubgridBatch
    from mrhyde_amd.abi import Problem, subgrid_workload

    n_axis, k, g = subgrid_workload()
    out = {"workload": "4096 unit-cube Q1 thermal subproblems, n^3 elements, n in {4,6,8,12,16}, "
                       "lognormal k (sigma 1), random Dirichlet data, Jacobi-PCG to 1e-10",
           "n_instances": int(len(n_axis)), "total_nodes": int(((n_axis.astype(np.int64) + 1) ** 3).sum())}
    nodes = (n_axis.astype(np.int64) + 1) ** 3
    for prec in ("fp64", "fp32-refined"):
        b = SubgridBatch(n_axis, k, g, tol=1e-10, precision=prec, use_graph=True)
        b.solve()
        ms = [b.solve() for _ in range(steps)]
        _, it, rel, refs = b.results()
        t = float(np.mean(ms))
        # per node and PCG iteration: 27 p + 8 k LDS reads, 8x(8+1) FMAs in the
        # stencil, ~10 flop of vector updates
        node_its = float((nodes * it).sum())
        out[prec] = {"ms_per_batch": round(t, 3), "solves_per_s": round(len(n_axis) / (t * 1e-3), 1),
                     "max_relres": float(rel.max()), "mean_iters": float(it.mean()), "wg_size": b.wg_size,
                     "mean_refinements": float(refs.mean()),
                     "node_iterations_per_s": round(node_its / (t * 1e-3), 1),
                     "est_tflops": round(154.0 * node_its / (t * 1e-3) / 1e12, 2),
                     "est_lds_TBps": round((35.0 * (8 if prec == "fp64" else 4)) * node_its / (t * 1e-3) / 1e12, 2)}
        b.close()
    # plain per-call path: one engine context + J/R assembly + Jacobi-CG per instance
    ko = np.concatenate([[0], np.cumsum(n_axis.astype(np.int64) ** 3)])
    xo = np.concatenate([[0], np.cumsum((n_axis.astype(np.int64) + 1) ** 3)])
    idx = list(range(0, len(n_axis), max(1, len(n_axis) // per_call_sample)))[:per_call_sample]
    t_all = t_solve = 0.0
    for i in idx:
        n = int(n_axis[i])
        t0 = time.perf_counter()
        ctx = Mhx(Problem(dim=3, order=1, quadrature=2, n=(n, n, n), k0=k[ko[i]:ko[i + 1]],
                          dbc_values=g[xo[i]:xo[i + 1]]))
        t1 = time.perf_counter()
        u = ctx.setDirichlet(np.zeros(ctx.n_dof))
        res, jac = ctx.assembleJacRes(u)
        du, _ = ctx.linearSolver(jac, res, solver="cg", tol=1e-10, max_iters=20000)
        t2 = time.perf_counter()
        t_all += t2 - t0
        t_solve += t2 - t1
        ctx.close()
    per = t_all / len(idx)
    per_solve = t_solve / len(idx)
    out["per_call"] = {"ms_per_instance": round(per * 1e3, 3), "solves_per_s": round(1.0 / per, 1),
                       "ms_assemble_solve_only": round(per_solve * 1e3, 3),
                       "solves_per_s_assemble_solve_only": round(1.0 / per_solve, 1),
                       "sample": f"{len(idx)} instances spread over the batch; one context per instance "
                                 "(set-up, J+R assembly, Jacobi-CG to 1e-10 with host buffers)"}
    # against the per-call path without its context set-up (the conservative ratio)
    base = out["per_call"]["solves_per_s_assemble_solve_only"]
    out["speedup_fp64_vs_per_call"] = round(out["fp64"]["solves_per_s"] / base, 1)
    out["speedup_fp32r_vs_per_call"] = round(out["fp32-refined"]["solves_per_s"] / base, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=96, help="elements per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--colouring", type=int, default=0, help="0 element colouring, 1 block colouring")
    ap.add_argument("--probe-gib", type=float, default=1.0, help="size of the plain-copy bandwidth probe")
    ap.add_argument("--skip-subgrid", action="store_true", help="skip the batched-subgrid measurement")
    ap.add_argument("--krylov-iters", type=int, default=100, help="GMRES iterations timed (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(each GPU would own a slab of elements; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from mrhyde_amd import Mhx
    from mrhyde_amd.abi import MHX_BUF_U, rng_fill

    n = args.n
    prob = c2_problem(n)
    t0 = time.perf_counter()
    g = Mhx(prob)
    g.set_option("assembly.colouring", args.colouring)
    t_setup = time.perf_counter() - t0
    u = rng_fill(7, g.n_dof, 0, 0.0, 1.0)
    g.setDirichlet(u)
    g.state_upload(MHX_BUF_U, u)
    print(f"[bench] setup {t_setup:.2f}s n_dof={g.n_dof} nnz={g.nnz} n_elem={g.n_elem}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        g.assemble_resident()
    g.synchronize()
    g.profile(True)
    g.profile_reset()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.assemble_resident()
    g.synchronize()
    dt = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    ms_step = dt * 1e3 / args.steps
    mdofs = g.n_dof / (ms_step * 1e-3) / 1e6

    kname = "k_thermal<jacres>"
    k_ms, k_cnt = prof.get(kname, (float("nan"), 0))
    launches_per_step = k_cnt / args.steps
    mean_launch_ms = k_ms / max(k_cnt, 1)
    n_vert = (n + 1) ** 3
    B = b_asm(g.n_elem, 27, 2, n_vert, 3)
    B_launch = B / launches_per_step
    achieved = B_launch / (mean_launch_ms * 1e-3) / 1e9
    B_min = 8 * g.nnz + 16 * g.n_dof + 8 * 2 * g.n_elem + 8 * 3 * n_vert
    pmc = PMC_TRAFFIC.get((n, args.colouring))
    traffic = (pmc["fetch_kib"] * pmc["fetch_scale"] + pmc["write_kib"] * pmc["write_scale"]) if pmc else None

    # achievable-bandwidth reference: plain device copy of known size
    import ctypes as C

    from mrhyde_amd import capi

    probe = None
    if args.probe_gib > 0:
        nd = int(args.probe_gib * 2 ** 30 / 8)
        ms = C.c_double()
        if capi.lib.mhx_copy_probe(nd, 5, C.byref(ms)) == 0:
            probe = round(16.0 * nd / (ms.value * 1e-3) / 1e9, 1)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n)

    krylov = krylov_bench(g, args.krylov_iters) if args.krylov_iters > 0 else None
    subgrid = None if args.skip_subgrid else subgrid_bench(max(3, min(args.steps, 10)))

    out = {
        "metric": "assembly MDOF/s (% HBM peak); subgrid solves/s; p50/p99 solve latency @1e-10",
        "value": round(mdofs, 2),
        "unit": "MDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal k0, uniform source, random u)",
        "config": {"workload": f"nonlinear thermal J+R assembly, {n}^3 Q2 hex, 3x3x3 Gauss, k=k0(1+T^2)",
                   "n_dof": g.n_dof, "nnz": g.nnz, "n_elem": g.n_elem, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, copy-kernel calibrated" if pmc else None,
                     "kernel": kname, "launches_per_step": launches_per_step,
                     "mean_launch_ms": round(mean_launch_ms, 5), "bytes_per_launch": B_launch,
                     "B_asm": B, "B_min": B_min, "copy_probe_GBps": probe,
                     "colouring": "block" if args.colouring else "element"},
        "cpu_baseline": cpu,
        "phases_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "krylov": krylov,
        "subgrid": subgrid,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
