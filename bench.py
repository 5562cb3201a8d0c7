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
baseline (the oracle on the full problem, OpenMP on the host's threads, with
one GMRES iteration); beside them the GMRES iteration rate on the assembled
Jacobian and the batched-subgrid solve rate.
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

# HBM traffic per launch of the dominant assembly kernel (96^3 Q2), keyed by
# (elements per axis, assembly.colouring), from rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE runs of this bench (mean of the 8 colour launches), scaled by the
# factors that make the same counters read the known size of the plain copy
# kernel: FETCH x 2048 B, WRITE x 1024 B.  Colouring 0: k_thermal<jacres>
# (profiles/k_thermal_q2hex_96cubed.md); colouring 3: k_q2box<jacres>, x-segments
# of 12 elements (profiles/k_q2box_q2hex_96cubed.md).  Only valid for those
# workloads.
PMC_TRAFFIC = {(96, 0): {"fetch_kib": 529028.5, "write_kib": 1032647.0, "fetch_scale": 2048, "write_scale": 1024},
               (96, 3): {"fetch_kib": 326291.6, "write_kib": 720445.1, "fetch_scale": 2048, "write_scale": 1024,
                         "segment": 12}}


def c2_problem(n: int):
    from mrhyde_amd.workloads import thermal_q2

    return thermal_q2(n)


def b_asm(ne: int, n_e: int, n_c: int, n_vert: int, dim: int) -> float:
    """Algorithmic bytes of one coloured J+R assembly (element-matrix read+write,
    u gather, residual read+write, coefficients, vertex coordinates)."""
    return ne * (16 * n_e * n_e + 24 * n_e + 8 * n_c) + 8 * dim * n_vert


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(n: int, gmres_iters: int = 3):
    """The oracle, which follows the reference's algorithm (worksets of 100
    elements, dense SFad<32> element duals, per-row column search, atomic
    scatter), on the full n^3 problem with OpenMP over worksets on every
    thread the process may use (OMP_NUM_THREADS, else the CPU count), and
    `gmres_iters` iterations of its GMRES(50) + Jacobi on the assembled
    Jacobian with OpenMP sparse products and vector updates."""
    from mrhyde_amd.abi import MHX_KRYLOV_GMRES, KrylovOpts, rng_fill
    from oracle import Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    o = Oracle(c2_problem(n))
    t_setup = time.perf_counter() - t0
    threads = o.set_threads(threads)
    u = rng_fill(7, o.n_dof, 0, 0.0, 1.0)
    o.setDirichlet(u)
    t0 = time.perf_counter()
    r, jac = o.assembleJacRes(u)
    dt = time.perf_counter() - t0
    opts = KrylovOpts()
    opts.solver, opts.restart, opts.max_iters, opts.tol = MHX_KRYLOV_GMRES, 50, gmres_iters, 1e-300
    t0 = time.perf_counter()
    o.linearSolver(jac, r, opts=opts)
    dt_g = time.perf_counter() - t0
    out = {"value": o.n_dof / dt / 1e6, "unit": "MDOF/s", "cores": threads, "kind": "port",
           "sample": f"oracle J+R assembly of the full {n}^3 Q2 problem ({o.n_elem} elements), OpenMP over "
                     f"worksets of 100, atomic scatter",
           "assembly_s": round(dt, 3), "setup_s": round(t_setup, 2),
           "gmres_iteration_ms": round(dt_g * 1e3 / gmres_iters, 2),
           "gmres": f"GMRES(50)+Jacobi, {gmres_iters} iterations from x = 0 (sparse products and vector updates "
                    f"on {threads} threads)",
           "nproc": os.cpu_count(), "cpu_model": cpu_model(), "threads": threads}
    del o, jac, r
    return out


def krylov_bench(g, iters: int):
    """GMRES(50) with Jacobi on the assembled Jacobian (device-resident), a
    fixed number of iterations: all-FP64, and FP32 Jacobian / vectors / inner
    GMRES inside FP64 refinement.  Per-iteration HIP-event time against the
    algorithmic bytes of an average GMRES(m) iteration,
    B_spmv + w N (m + 3) with w the vector width; B_spmv counts the values,
    row pointers and vectors (the columns come from the closed form of the
    lattice pattern, no index stream is read)."""
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
        b_it = wa * nnz + 4 * (N + 1) + 2 * wv * N + wv * N * (m + 3)
        out[name] = {"iterations": st["iters"], "refinements": st["refinements"], "rel_res": st["rel_res"],
                     "wall_s": round(dt, 3), "ms_per_iteration": round(per_it, 4),
                     "bytes_per_iteration": b_it, "achieved_GBps": round(b_it / (per_it * 1e-3) / 1e9, 1)}
    # time to solution at 1e-10 with the smoothed-aggregation preconditioner
    # (hierarchy built from this Jacobian inside the timed solve)
    g.state_upload(MHX_BUF_DU, np.zeros(N))
    g.synchronize()
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    st = g.solve_resident(solver="gmres", restart=m, tol=1e-10, max_iters=1000, precond="amg")
    g.synchronize()
    dt = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    setup_ms = prof.get("LinearAlgebraInterface::buildAMGPreconditioner", (0.0, 0))[0]
    vc_ms, vc_n = prof.get("MueLu::Hierarchy::Iterate", (0.0, 0))
    out["amg_fp64_to_1e-10"] = {"solver": "GMRES(50) + smoothed-aggregation AMG, Chebyshev(2) V-cycle",
                                "iterations": st["iters"], "rel_res": st["rel_res"], "solve_latency_ms": round(dt * 1e3, 1),
                                "setup_ms": round(setup_ms, 1), "vcycle_ms": round(vc_ms / max(vc_n, 1), 3),
                                "levels": [g.amg_level(lv, "A")[1]["rows"]
                                           for lv in range(g.amg_level(0, "A")[1]["levels"])]}
    # the same with the FP32 hierarchy and inner GMRES in FP64 refinement
    x64 = g.state_download(MHX_BUF_DU)
    g.state_upload(MHX_BUF_DU, np.zeros(N))
    g.synchronize()
    t0 = time.perf_counter()
    st = g.solve_resident(solver="gmres", restart=m, tol=1e-10, max_iters=1000, precond="amg-fp32")
    g.synchronize()
    dt = time.perf_counter() - t0
    x32 = g.state_download(MHX_BUF_DU)
    out["amg_fp32_refined_to_1e-10"] = {"iterations": st["iters"], "refinements": st["refinements"],
                                        "rel_res": st["rel_res"], "solve_latency_ms": round(dt * 1e3, 1),
                                        "max_abs_diff_vs_fp64": float(np.abs(x32 - x64).max()),
                                        "max_abs_fp64": float(np.abs(x64).max())}
    return out


def ns_bench(n: int, steps: int, newton: bool = True):
    """Steady Navier-Stokes, Q1-Q1 with SUPG/PSPG: lid-driven cavity on [0,1]^3
    with n^3 hexes, Re = 100 (viscosity 0.01, lid speed 1 on y = 1, no-slip
    elsewhere, pressure pinned at the origin), initial velocity N(0, 1e-3)
    seeded.  J+R assembly rate and its roofline; the first Newton system
    solved to 1e-10 with nodal block Jacobi and with the AMG cycle smoothed
    by nodal block Jacobi; then Newton with GMRES(50) + that AMG cycle to
    1e-10 per step, FP32 hierarchy and inner solve in FP64 refinement,
    against the all-FP64 Newton (difference stated)."""
    from mrhyde_amd import Mhx
    from mrhyde_amd.abi import MHX_BUF_U
    from mrhyde_amd.workloads import cavity, cavity_state

    g = Mhx(cavity(n))
    u0 = g.setDirichlet(cavity_state(g.n_dof))
    g.state_upload(MHX_BUF_U, u0)
    for _ in range(2):
        g.assemble_resident()
    g.synchronize()
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.assemble_resident()
    g.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = g.profile_read()
    g.profile(False)
    k_ms, k_cnt = prof.get("k_ns<jacres>", (float("nan"), 0))
    per_launch = k_ms / max(k_cnt, 1)
    n_vert = (n + 1) ** 3
    B = b_asm(g.n_elem, 32, 0, n_vert, 3)
    lpl = k_cnt / steps
    out = {"workload": f"3-D lid-driven cavity Re=100, {n}^3 Q1-Q1 hexes, SUPG+PSPG, 4 unknowns per node",
           "n_dof": g.n_dof, "nnz": g.nnz, "assembly_ms": round(dt * 1e3, 4),
           "assembly_MDOFps": round(g.n_dof / dt / 1e6, 2),
           "roofline": {"kernel": "k_ns<jacres>", "B_asm": B, "launches_per_step": lpl,
                        "mean_launch_ms": round(per_launch, 5),
                        "achieved_GBps": round(B / lpl / (per_launch * 1e-3) / 1e9, 1),
                        "frac": round(B / lpl / (per_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    # the first Newton system (J, res at u0, resident) solved to 1e-10 from du = 0
    from mrhyde_amd.abi import MHX_BUF_DU

    lin = {}
    for name, sm, prec, m in (("block_jacobi", 0, "block jacobi", 150), ("amg_block_jacobi_smoother", 1, "amg", 150)):
        g.set_option("amg.smoother", sm)
        g.state_upload(MHX_BUF_DU, np.zeros(g.n_dof))
        g.synchronize()
        t0 = time.perf_counter()
        st = g.solve_resident(solver="gmres", restart=m, tol=1e-10, max_iters=3000, precond=prec)
        g.synchronize()
        lin[name] = {"solver": f"GMRES({m})", "iterations": st["iters"], "rel_res": st["rel_res"],
                     "wall_s": round(time.perf_counter() - t0, 3)}
    g.set_option("amg.smoother", -1)
    out["first_newton_system"] = lin
    if newton:
        kw = dict(max_iters=10, nl_tol=1e-8, solver="gmres", restart=50, tol=1e-10, linear_max_iters=2000)
        res = {}
        g.set_option("amg.smoother", 1)
        for name, prec in (("fp32-refined", "amg-fp32"), ("fp64", "amg")):
            t0 = time.perf_counter()
            u, st = g.nonlinearSolver(u0.copy(), precond=prec, **kw)
            res[name] = (u, st, time.perf_counter() - t0)
        g.set_option("amg.smoother", -1)
        u32, s32, t32 = res["fp32-refined"]
        u64, s64, t64 = res["fp64"]
        out["newton"] = {
            name: {"wall_s": round(t, 3), "nl_iters": st["nl_iters"], "krylov_iters": st["iters"],
                   "refinements": st["refinements"], "res_norm0": st["res_norm0"], "res_norm": st["res_norm"],
                   "last_linear_rel_res": st["rel_res"]}
            for name, (u, st, t) in res.items()}
        out["newton"]["max_abs_diff_fp32r_vs_fp64"] = float(np.abs(u32 - u64).max())
        out["newton"]["max_abs_u"] = float(np.abs(u64).max())
        out["newton"]["preconditioner"] = (
            "GMRES(50) + AMG with nodal block-Jacobi smoothing (deviation: the configuration prescribes FP32 "
            "block Jacobi, which stalls near 3e-8 on this cavity; see first_newton_system.block_jacobi)")
    g.close()
    return out


def c2_newton_bench(n: int):
    """The nonlinear Q2 thermal problem solved end to end: Newton from u = 0
    ("max nonlinear iters" 5, no backtracking), each linear solve GMRES(50) to
    1e-10 with the smoothed-aggregation preconditioner (rebuilt per step, as
    the reference's MueLu set-up with reuse "none").  Wall time from host u to
    host u, per-phase times from the engine's timers."""
    from mrhyde_amd import Mhx

    g = Mhx(c2_problem(n))
    u = g.setDirichlet(np.zeros(g.n_dof))
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    u, st = g.nonlinearSolver(u, max_iters=5, nl_tol=1e-12, backtracking=False, solver="gmres", restart=50,
                              tol=1e-10, linear_max_iters=1000, precond="amg")
    wall = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    # the same with the FP32 hierarchy and inner GMRES in FP64 refinement
    u32 = g.setDirichlet(np.zeros(g.n_dof))
    t0 = time.perf_counter()
    u32, st32 = g.nonlinearSolver(u32, max_iters=5, nl_tol=1e-12, backtracking=False, solver="gmres", restart=50,
                                  tol=1e-10, linear_max_iters=1000, precond="amg-fp32")
    wall32 = time.perf_counter() - t0
    # hierarchy reuse across the Newton steps (coarse levels of the first step
    # kept, fine-level smoother data refreshed), FP64 and FP32-refined
    g.set_option("amg.reuse", 1)
    reuse = {}
    for name, prec in (("fp64", "amg"), ("fp32_refined", "amg-fp32")):
        ur = g.setDirichlet(np.zeros(g.n_dof))
        t0 = time.perf_counter()
        ur, str_ = g.nonlinearSolver(ur, max_iters=5, nl_tol=1e-12, backtracking=False, solver="gmres", restart=50,
                                     tol=1e-10, linear_max_iters=1000, precond=prec)
        reuse[name] = {"wall_s": round(time.perf_counter() - t0, 3), "krylov_iters": str_["iters"],
                       "res_norm": str_["res_norm"], "last_linear_rel_res": str_["rel_res"],
                       "max_abs_diff_vs_fp64": float(np.abs(ur - u).max())}
    g.set_option("amg.reuse", 0)
    # the configuration's own preconditioner: GMRES(50) + Jacobi to 1e-10 per
    # step (the AMG legs above are the faster substitute)
    uj = g.setDirichlet(np.zeros(g.n_dof))
    t0 = time.perf_counter()
    uj, stj = g.nonlinearSolver(uj, max_iters=5, nl_tol=1e-12, backtracking=False, solver="gmres", restart=50,
                                tol=1e-10, linear_max_iters=20000, precond="jacobi")
    jacobi = {"preconditioner": "Jacobi (the configuration's)", "wall_s": round(time.perf_counter() - t0, 3),
              "nl_iters": stj["nl_iters"], "krylov_iters": stj["iters"], "res_norm": stj["res_norm"],
              "last_linear_rel_res": stj["rel_res"], "max_abs_diff_vs_amg": float(np.abs(uj - u).max())}
    g.close()
    ph = {k: round(v[0], 1) for k, v in prof.items()
          if k in ("AssemblyManager::assembleJacRes", "LinearAlgebraInterface::linearSolver",
                   "LinearAlgebraInterface::buildAMGPreconditioner", "MueLu::Hierarchy::Iterate")}
    return {"workload": f"Newton on the {n}^3 Q2 nonlinear thermal problem from u = 0, 5 steps max, "
                        "GMRES(50) + AMG to 1e-10 per step",
            "preconditioner": "smoothed-aggregation AMG (deviation: the configuration prescribes Jacobi; "
                              "the Jacobi run is reported under 'config_jacobi')",
            "config_jacobi": jacobi,
            "wall_s": round(wall, 3), "nl_iters": st["nl_iters"], "krylov_iters": st["iters"],
            "res_norm0": st["res_norm0"], "res_norm": st["res_norm"], "last_linear_rel_res": st["rel_res"],
            "phases_ms": ph,
            "fp32_refined": {"wall_s": round(wall32, 3), "nl_iters": st32["nl_iters"], "krylov_iters": st32["iters"],
                             "refinements": st32["refinements"], "res_norm": st32["res_norm"],
                             "last_linear_rel_res": st32["rel_res"],
                             "max_abs_diff_vs_fp64": float(np.abs(u32 - u).max()),
                             "max_abs_u": float(np.abs(u).max())},
            "amg_reuse": reuse}


def c5_requests(count: int, seed: int):
    from mrhyde_amd.workloads import serve_requests

    return serve_requests(count, seed)


CACHED = 256  # idle contexts a server keeps (within its HBM budget)
LDS_PEAK_TBS = 78.6  # MI355X LDS: 128 B / clock / CU x 256 CUs x 2.4 GHz
FP64_MFMA_PEAK_TFS = 78.6  # MI355X dense FP64 matrix peak


def serving_bench(n_req: int = 240, rate: float = 200.0, seed: int = 5, per_call_sample: int = 12):
    """Poisson arrivals at `rate` of C5 requests into the server (8 concurrent
    requests, resident contexts reused, up to CACHED idle contexts kept),
    Jacobi-CG to 1e-10; latency from submit to result.  Then the same
    requests all at once (capacity) in FP64 and with the FP32-inner /
    FP64-refined CG, the AMG-preconditioned forms, Poisson arrivals at 70 %
    of the mixed-precision AMG server's capacity, and the plain per-call
    path (a fresh context per request, sequential) on a sample."""
    from mrhyde_amd import Mhx, Server

    probs = c5_requests(n_req, seed)
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = {"workload": f"{n_req} thermal solve requests (2-D Q1 32^2..256^2, 3-D Q1/Q2 16^3..64^3, log-uniform), "
                       f"Jacobi-CG to 1e-10", "arrival_rate": rate,
           "total_dof": int(sum(p.n_dof for p in probs))}

    def run(srv, timed_arrivals):
        t0 = time.perf_counter()
        tickets = []
        for p, ta in zip(probs, timed_arrivals):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(p))
        res = [srv.poll(t) for t in tickets]
        wall = time.perf_counter() - t0
        lat = np.array([r[2] for r in res])
        rel = max(r[1]["rel_res"] for r in res)
        return {"wall_s": round(wall, 3), "solves_per_s": round(n_req / wall, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "max_linear_rel_res": rel, "server": srv.info()}

    srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond="jacobi", tol=1e-10)
    out["poisson"] = run(srv, arrivals)
    srv.close()
    for name, prec in (("saturated_fp64", "jacobi"), ("saturated_fp32_refined", "jacobi-fp32")):
        srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond=prec, tol=1e-10)
        out[name] = run(srv, np.zeros(n_req))
        srv.close()
    # the same arrivals with the smoothed-aggregation V-cycle as the CG
    # preconditioner (the reference's default preconditioner family)
    srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond="amg", tol=1e-10)
    out["poisson_amg"] = run(srv, arrivals)
    srv.close()
    srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond="amg", tol=1e-10)
    out["saturated_amg"] = run(srv, np.zeros(n_req))
    srv.close()
    # mixed precision: FP32 hierarchy and inner CG, FP64 residual and refinement
    srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond="amg-fp32", tol=1e-10)
    out["saturated_amg_fp32_refined"] = run(srv, np.zeros(n_req))
    srv.close()
    # Poisson arrivals at a rate the mixed-precision AMG server sustains
    # (70 % of its saturated throughput): latency without queue growth
    lam = 0.7 * out["saturated_amg_fp32_refined"]["solves_per_s"]
    arr = np.cumsum(np.random.default_rng(seed + 2).exponential(1.0 / lam, n_req))
    srv = Server(max_active=8, max_cached=CACHED, solver="cg", precond="amg-fp32", tol=1e-10)
    out["poisson_sustained_amg_fp32_refined"] = dict(run(srv, arr), arrival_rate=round(lam, 1))
    srv.close()
    idx = np.linspace(0, n_req - 1, per_call_sample).astype(int)
    t_all = 0.0
    for i in idx:
        t0 = time.perf_counter()
        g = Mhx(probs[i])
        u = g.setDirichlet(np.zeros(g.n_dof))
        g.nonlinearSolver(u, max_iters=2, nl_tol=1e-8, solver="cg", tol=1e-10, linear_max_iters=20000)
        g.close()
        t_all += time.perf_counter() - t0
    # per-call time scaled to the whole request set by DOF count
    dof_sample = sum(probs[i].n_dof for i in idx)
    per_call_wall = t_all * out["total_dof"] / dof_sample
    out["per_call"] = {"sample": f"{per_call_sample} requests evenly spaced over the set, fresh context each, "
                                 "sequential; time scaled to the whole set by DOF count",
                       "solves_per_s": round(n_req / per_call_wall, 1)}
    best = max(out["saturated_fp64"]["solves_per_s"], out["saturated_fp32_refined"]["solves_per_s"])
    out["speedup_served_vs_per_call"] = round(best / out["per_call"]["solves_per_s"], 2)
    out["speedup_amg_vs_jacobi_served"] = round(out["saturated_amg"]["solves_per_s"] /
                                                out["saturated_fp64"]["solves_per_s"], 2)
    # the served mixed-precision AMG form against the plain per-call path
    # (Jacobi-CG, fresh context per request), both to 1e-10
    out["speedup_served_amg_fp32r_vs_per_call"] = round(out["saturated_amg_fp32_refined"]["solves_per_s"] /
                                                        out["per_call"]["solves_per_s"], 2)
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
        # the two ceilings of the on-chip solve: vector FP64 (FP32 inner:
        # the FP32 vector rate) and LDS bandwidth (128 B per clock per CU,
        # 256 CUs at 2.4 GHz = 78.6 TB/s)
        flop_peak = 78.6 if prec == "fp64" else 157.3
        out[prec]["roofline"] = {"bound": "lds", "achieved": out[prec]["est_lds_TBps"], "peak": LDS_PEAK_TBS,
                                 "unit": "TB/s", "frac": round(out[prec]["est_lds_TBps"] / LDS_PEAK_TBS, 4),
                                 "flop_frac": round(out[prec]["est_tflops"] / flop_peak, 4),
                                 "flop_peak_TFLOPs": flop_peak}
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
    ap.add_argument("--colouring", type=int, default=-1,
                    help="assembly schedule: -1 the context's default (3, pencil streaming, on this box mesh), "
                         "0 element colouring, 1 block colouring, 2 x-segments")
    ap.add_argument("--segment", type=int, default=0, help="pencil streaming: elements per x-segment (0 = default)")
    ap.add_argument("--probe-gib", type=float, default=1.0, help="size of the plain-copy bandwidth probe")
    ap.add_argument("--skip-subgrid", action="store_true", help="skip the batched-subgrid measurement")
    ap.add_argument("--krylov-iters", type=int, default=100, help="GMRES iterations timed (0 = skip)")
    ap.add_argument("--no-newton", action="store_true", help="skip the end-to-end Newton solve")
    ap.add_argument("--ns-n", type=int, default=64, help="Navier-Stokes cavity elements per axis (0 = skip)")
    ap.add_argument("--ns-no-newton", action="store_true", help="Navier-Stokes: assembly only")
    ap.add_argument("--serve-requests", type=int, default=240, help="serving requests (0 = skip)")
    ap.add_argument("--mfma-probe", type=int, default=1, help="time the matrix-core element contraction (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the multi-GPU form "
              "(each GPU owning a z-slab of elements) is not built", file=sys.stderr)
        sys.exit(2)

    from mrhyde_amd import Mhx
    from mrhyde_amd.abi import MHX_BUF_U, rng_fill

    n = args.n
    prob = c2_problem(n)
    t0 = time.perf_counter()
    g = Mhx(prob)
    if args.colouring >= 0:
        g.set_option("assembly.colouring", args.colouring)
    if args.segment > 0:
        g.set_option("assembly.segment", args.segment)
    colouring = g.get_option("assembly.colouring")
    segment = g.get_option("assembly.segment")
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

    kname = "k_q2box<jacres>" if colouring == 3 else "k_thermal<jacres>"
    k_ms, k_cnt = prof.get(kname, (float("nan"), 0))
    launches_per_step = k_cnt / args.steps
    mean_launch_ms = k_ms / max(k_cnt, 1)
    n_vert = (n + 1) ** 3
    B = b_asm(g.n_elem, 27, 2, n_vert, 3)
    B_launch = B / launches_per_step
    achieved = B_launch / (mean_launch_ms * 1e-3) / 1e9
    B_min = 8 * g.nnz + 16 * g.n_dof + 8 * 2 * g.n_elem + 8 * 3 * n_vert
    pmc = PMC_TRAFFIC.get((n, colouring))
    if pmc and pmc.get("segment", segment) != segment:
        pmc = None  # measured at another segment length
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

    # the dense matrix-core form of the element contraction, on as many
    # elements as one colour launch of this assembly
    mfma = None
    if args.mfma_probe:
        ne = int(round(g.n_elem / launches_per_step))
        mfma = {"elements": ne, "mfma_per_element": 108, "assembly_launch_ms": round(mean_launch_ms, 4),
                "note": "dense B^T D B contraction only (no point physics, gather or scatter) against "
                        "the whole sum-factorised assembly launch"}
        for variant, name in ((0, "table_in_lds"), (1, "table_in_registers")):
            ms, err = C.c_double(), C.c_double()
            if capi.lib.mhx_mfma_probe(ne, 10, variant, C.byref(ms), C.byref(err)) == 0:
                tflops = 108 * 2048.0 * ne / (ms.value * 1e-3) / 1e12
                mfma[name] = {"ms_per_pass": round(ms.value, 4), "TFLOPs": round(tflops, 2),
                              "mfma_frac": round(tflops / FP64_MFMA_PEAK_TFS, 4), "max_rel_err": err.value}

    def step(name, fn):
        t0 = time.perf_counter()
        print(f"[bench] {name} ...", file=sys.stderr, flush=True)
        r = fn()
        print(f"[bench] {name} done in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        return r

    print(f"[bench] assembly {ms_step:.3f} ms/step, {mdofs:.1f} MDOF/s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu_baseline else step("cpu baseline", lambda: cpu_baseline(n))
    krylov = step("krylov", lambda: krylov_bench(g, args.krylov_iters)) if args.krylov_iters > 0 else None
    g.close()
    newton = None if args.no_newton else step("newton", lambda: c2_newton_bench(n))
    steps_small = max(3, min(args.steps, 10))
    subgrid = None if args.skip_subgrid else step("subgrid", lambda: subgrid_bench(steps_small))
    ns = step("navier stokes", lambda: ns_bench(args.ns_n, steps_small, not args.ns_no_newton)) if args.ns_n > 0 else None
    serving = step("serving", lambda: serving_bench(args.serve_requests)) if args.serve_requests > 0 else None

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
                     "traffic_per": "launch",
                     "kernel": kname, "launches_per_step": launches_per_step,
                     "mean_launch_ms": round(mean_launch_ms, 5), "bytes_per_launch": B_launch,
                     "B_asm": B, "B_min": B_min, "copy_probe_GBps": probe,
                     "colouring": {0: "element", 1: "block", 2: "x-segments", 3: "pencil streaming"}[colouring],
                     "segment": segment if colouring == 3 else None},
        "cpu_baseline": cpu,
        "mfma_contraction_probe": mfma,
        "phases_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "krylov": krylov,
        "newton": newton,
        "subgrid": subgrid,
        "navier_stokes": ns,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
