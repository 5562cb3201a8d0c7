"""GPU parity: libmhx.so (HIP, gfx950) against the CPU oracle on the same
seeded inputs.  Bars: CRS pattern bit-exact; residual and Jacobian entries
within 1e-12 relative (plus 1e-14 of the row's largest entry for entries that
cancel to ~0); Krylov solutions at <= 1e-10 true relative residual."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_entries_close, spmv_np, to_scipy
from mrhyde_amd.abi import (MHX_KRYLOV_CG, MHX_KRYLOV_GMRES, MHX_PHYS_NAVIER_STOKES, MHX_PREC_JACOBI, KrylovOpts,
                            NewtonOpts, Problem)

pytestmark = pytest.mark.gpu
PI = np.pi


def lognormal(n, s, seed):
    return np.exp(s * np.random.default_rng(seed).standard_normal(n))


def make_configs():
    cfg = {}
    cfg["2d_q1_poisson_c1shape"] = Problem(dim=2, order=1, quadrature=2, n=(16, 16),
                                           modes=[(2 * PI * PI, (PI, PI))])
    cfg["2d_q2_nonlinear_perturbed"] = Problem(dim=2, order=2, quadrature=4, n=(7, 5), diff_u2=1.0,
                                               k0=lognormal(35, 0.5, 1), modes=[(3.0, (PI, 2 * PI))],
                                               perturb=0.3, perturb_seed=2, dbc_sides=1 | 2)
    cfg["2d_q4_quad8"] = Problem(dim=2, order=4, quadrature=8, n=(5, 4), modes=[(8 * PI * PI, (2 * PI, 2 * PI))])
    cfg["2d_q3_ragged_1x7"] = Problem(dim=2, order=3, quadrature=6, n=(1, 7), k0=lognormal(7, 1.0, 3),
                                      src_elem=np.linspace(-1, 1, 7), dbc_sides=4)
    cfg["2d_q1_underintegrated"] = Problem(dim=2, order=1, quadrature=0, n=(5, 6), src_elem=np.ones(30))
    # the highest built 2-D orders: four elements per workgroup, so the x-segment
    # schedule runs segments of 4, 4 and 1 (Q5) and 4 and 2 (Q6) elements
    cfg["2d_q5_quad10_perturbed"] = Problem(dim=2, order=5, quadrature=10, n=(9, 3), diff_u2=0.5,
                                            k0=lognormal(27, 0.5, 13), modes=[(5.0, (PI, 2 * PI))], perturb=0.15,
                                            perturb_seed=14, dbc_sides=1 | 4)
    cfg["2d_q6_quad12"] = Problem(dim=2, order=6, quadrature=12, n=(6, 2), diff_u2=1.0, k0=lognormal(12, 0.5, 15),
                                  src_elem=np.linspace(0, 1, 12), dbc_sides=1 | 8)
    cfg["3d_q1_lognormal"] = Problem(dim=3, order=1, quadrature=2, n=(6, 5, 4), k0=lognormal(120, 1.0, 4),
                                     src_elem=np.random.default_rng(5).uniform(0, 1, 120))
    cfg["3d_q1_single_element"] = Problem(dim=3, order=1, quadrature=2, n=(1, 1, 1), dbc_sides=0,
                                          src_elem=np.ones(1), diff_u2=1.0)
    cfg["3d_q2_c2shape"] = Problem(dim=3, order=2, quadrature=4, n=(5, 4, 3), diff_u2=1.0,
                                   k0=lognormal(60, 0.5, 6), src_elem=np.random.default_rng(7).uniform(0, 1, 60),
                                   dbc_sides=1 | 2)
    cfg["3d_q2_perturbed_modes"] = Problem(dim=3, order=2, quadrature=4, n=(3, 3, 4), diff_u2=0.7,
                                           modes=[(2.0, (PI, PI, PI)), (0.5, (2 * PI, 0, 3 * PI))],
                                           perturb=0.25, perturb_seed=9)
    cfg["3d_q2_quad2"] = Problem(dim=3, order=2, quadrature=2, n=(3, 2, 2), k0=lognormal(12, 0.5, 8))
    # Neumann sides (thermal boundaryResidual)
    flux2 = [((0.7, -1.3), (PI, 2 * PI), (0.3, 0.0)), ((0.0, 2.0), (2 * PI, 0.0), (0.0, PI / 2))]
    cfg["2d_q2_neumann_perturbed"] = Problem(dim=2, order=2, quadrature=4, n=(6, 5), k0=lognormal(30, 0.5, 9),
                                             diff_u2=0.5, dbc_sides=1, neumann_sides=2 | 4 | 8, neumann_flux=flux2,
                                             perturb=0.2, perturb_seed=10)
    flux3 = [((0.5, 1.0, -2.0), (PI, PI, 2 * PI), (0.0, 0.4, 0.0))]
    cfg["3d_q2_neumann"] = Problem(dim=3, order=2, quadrature=4, n=(4, 3, 3), k0=lognormal(36, 0.5, 11),
                                   diff_u2=1.0, dbc_sides=1 | 2, neumann_sides=4 | 16 | 32, neumann_flux=flux3,
                                   perturb=0.2, perturb_seed=12)
    # navier stokes, Q1-Q1
    cfg["2d_ns_channel_pspg"] = Problem(dim=2, order=1, quadrature=2, n=(10, 4), hi=(5.0, 1.0),
                                        physics=MHX_PHYS_NAVIER_STOKES, dbc_sides=4 | 8, pspg=True,
                                        force=(1.0, 0.0))
    cfg["2d_ns_supg_perturbed"] = Problem(dim=2, order=1, quadrature=2, n=(6, 5), physics=MHX_PHYS_NAVIER_STOKES,
                                          rho=1.2, mu=0.02, supg=True, pspg=True, dbc_sides=1 | 4, lid_sides=8,
                                          lid_u=(1.0, 0.0), comp_sides=(0, 2), pin_pressure=True,
                                          force=(0.3, -0.1), perturb=0.25, perturb_seed=4)
    cfg["3d_ns_lid_cavity"] = Problem(dim=3, order=1, quadrature=2, n=(5, 4, 3), physics=MHX_PHYS_NAVIER_STOKES,
                                      mu=0.01, supg=True, pspg=True, dbc_sides=1 | 2 | 4 | 16 | 32, lid_sides=8,
                                      lid_u=(1.0, 0.0, 0.0), pin_pressure=True, perturb=0.2, perturb_seed=6)
    return cfg


CONFIGS = make_configs()


@pytest.fixture(scope="module")
def pairs():
    from mrhyde_amd import Mhx
    from oracle import Oracle

    return {k: (Mhx(p), Oracle(p)) for k, p in CONFIGS.items()}


def random_state(o, seed):
    u = np.random.default_rng(seed).standard_normal(o.n_dof)
    return o.setDirichlet(u)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_pattern_and_dirichlet_bit_exact(pairs, name):
    g, o = pairs[name]
    assert (g.n_dof, g.nnz) == (o.n_dof, o.nnz)
    rg, cg = g.crs_pattern()
    ro, co = o.crs_pattern()
    np.testing.assert_array_equal(rg, ro)
    np.testing.assert_array_equal(cg, co)
    mo, _ = o.dirichlet()
    np.testing.assert_array_equal(g.dirichlet_mask(), mo)
    np.testing.assert_allclose(g.dof_coords(), o.dof_coords(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("colouring", [2, 1, 0])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_assembly_parity(pairs, name, colouring):
    g, o = pairs[name]
    print(f"[parity] {name} colouring={colouring} n_dof={g.n_dof} nnz={g.nnz}", flush=True)
    g.set_option("assembly.colouring", colouring)
    u = random_state(o, 17)
    # stale values in the output buffers must not leak into an overwrite-mode assembly
    rg = np.full(g.n_dof, 7.0)
    jg = np.full(g.nnz, -3.0)
    g.assemble(u, rg, jg)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"{name} Jacobian")
    assert_entries_close(rg, ro, None, what=f"{name} residual")
    # residual-only pass (assembleRes) gives the same residual
    assert_entries_close(g.assembleRes(u), ro, None, what=f"{name} residual-only")
    g.set_option("assembly.colouring", 1)


@pytest.mark.parametrize("colouring", [2, 1, 0])
def test_accumulate_and_device_pointers(pairs, colouring):
    from mrhyde_amd.abi import MHX_BUF_JAC, MHX_BUF_RES, MHX_BUF_U

    g, o = pairs["3d_q2_c2shape"]
    g.set_option("assembly.colouring", colouring)
    u = random_state(o, 3)
    ro, jo = o.assembleJacRes(u)
    g.state_upload(MHX_BUF_U, u)
    g.assemble_resident()
    g.assemble_resident(accumulate=True)
    g.set_option("assembly.colouring", 1)
    rp, ci = o.crs_pattern()
    mask, _ = o.dirichlet()
    jo2 = 2 * jo
    diag = np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i) for i in np.where(mask)[0]])
    jo2[diag] = 1.0  # dofConstraints replaces the diagonal, it does not add
    assert_entries_close(g.state_download(MHX_BUF_JAC), jo2, rp, what="accumulated Jacobian")
    assert_entries_close(g.state_download(MHX_BUF_RES), 2 * ro, None, what="accumulated residual")
    # the resident buffers feed the device-pointer solve path
    x, st = g.linearSolver(jo, ro, solver="gmres", tol=1e-10, max_iters=3000)
    assert st["rel_res"] <= 1e-10


@pytest.mark.parametrize("name,solver", [("2d_q1_poisson_c1shape", MHX_KRYLOV_CG),
                                         ("3d_q1_lognormal", MHX_KRYLOV_CG),
                                         ("2d_q2_nonlinear_perturbed", MHX_KRYLOV_GMRES),
                                         ("3d_q2_c2shape", MHX_KRYLOV_GMRES),
                                         ("2d_q4_quad8", MHX_KRYLOV_GMRES)])
def test_linear_solve_parity(pairs, name, solver):
    import scipy.sparse.linalg as spla

    from mrhyde_amd import krylov_opts

    g, o = pairs[name]
    u = o.setDirichlet(np.zeros(o.n_dof))
    ro, jo = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    opts = krylov_opts(solver="cg" if solver == MHX_KRYLOV_CG else "gmres", tol=1e-10, max_iters=5000, restart=50)
    x, st = g.linearSolver(jo, ro, opts=opts)
    rel = np.linalg.norm(ro - spmv_np(rp, ci, jo, x)) / np.linalg.norm(ro)
    assert st["rel_res"] <= 1e-10 and rel <= 1.5e-10, (st, rel)
    x_direct = spla.spsolve(to_scipy(rp, ci, jo).tocsc(), ro)
    np.testing.assert_allclose(x, x_direct, rtol=0, atol=1e-7 * np.abs(x_direct).max())
    # the CRS product on the device
    y = g.apply(jo, x)
    np.testing.assert_allclose(y, spmv_np(rp, ci, jo, x), rtol=1e-13, atol=1e-14 * np.abs(y).max())


@pytest.mark.parametrize("name", ["3d_q2_c2shape", "3d_ns_lid_cavity", "2d_q1_poisson_c1shape"])
def test_spmv_lane_widths(pairs, name):
    """The CRS product at every lane count per row (rows of 9 to 125 entries:
    lanes with several four-entry passes, one pass, and only the tail loop)."""
    g, o = pairs[name]
    u = random_state(o, 11)
    _, jo = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    x = np.random.default_rng(12).standard_normal(o.n_dof)
    ref = spmv_np(rp, ci, jo, x)
    default = g.get_option("krylov.spmv_lanes")
    for lanes in (4, 8, 16, 32, 64):
        g.set_option("krylov.spmv_lanes", lanes)
        y = g.apply(jo, x)
        np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-14 * np.abs(ref).max(), err_msg=f"lanes={lanes}")
    g.set_option("krylov.spmv_lanes", default)
    from mrhyde_amd import MhxError

    with pytest.raises(MhxError):
        g.set_option("krylov.spmv_lanes", 12)


@pytest.mark.parametrize("name", ["3d_q2_c2shape", "3d_ns_lid_cavity", "2d_q1_poisson_c1shape", "2d_q4_quad8",
                                  "2d_ns_supg_perturbed", "2d_q3_ragged_1x7"])
def test_lattice_columns_match_colidx(pairs, name):
    """The sparse products that take their columns from the closed form of the
    lattice pattern visit the same entries in the same order as the ones that
    read colidx: products, and whole GMRES solves with Jacobi and with AMG
    (level-0 smoothing and residuals), are bit-identical."""
    from mrhyde_amd import krylov_opts

    g, o = pairs[name]
    # the cavity linearised at the start of a lid-driven run (its Jacobian at
    # u ~ N(0, 1) defeats every preconditioner here)
    scale = 1e-3 if "ns" in name else 1.0
    u = o.setDirichlet(scale * np.random.default_rng(13).standard_normal(o.n_dof))
    r, jo = o.assembleJacRes(u)
    x = np.random.default_rng(14).standard_normal(o.n_dof)
    assert g.get_option("krylov.lattice_columns") == 1
    out = {}
    for lat in (1, 0):
        g.set_option("krylov.lattice_columns", lat)
        res = [g.apply(jo, x)]
        for prec in ("jacobi", "amg", "amg-fp32"):
            xs, st = g.linearSolver(jo, r, opts=krylov_opts(solver="gmres", tol=1e-10, max_iters=60, precond=prec))
            res += [xs, np.array([st["iters"], st["rel_res"]])]
        out[lat] = res
    g.set_option("krylov.lattice_columns", 1)
    for a, b in zip(out[1], out[0]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name,solver", [("2d_q1_poisson_c1shape", "cg"), ("3d_q1_lognormal", "cg"),
                                         ("3d_q2_c2shape", "gmres"), ("2d_q4_quad8", "gmres")])
def test_mixed_precision_solve(pairs, name, solver):
    """FP32 Jacobian, preconditioner and inner Krylov inside FP64 iterative
    refinement: the FP64 true residual reaches 1e-10 and the result agrees
    with the all-FP64 solve."""
    from mrhyde_amd import krylov_opts

    g, o = pairs[name]
    u = o.setDirichlet(np.zeros(o.n_dof))
    ro, jo = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    print(f"[mixed] {name} {solver} n_dof={g.n_dof}", flush=True)
    x32, st = g.linearSolver(jo, ro, opts=krylov_opts(solver=solver, tol=1e-10, max_iters=5000, precond="jacobi-fp32"))
    x64, st64 = g.linearSolver(jo, ro, opts=krylov_opts(solver=solver, tol=1e-10, max_iters=5000))
    rel = np.linalg.norm(ro - spmv_np(rp, ci, jo, x32)) / np.linalg.norm(ro)
    assert st["rel_res"] <= 1e-10 and rel <= 1.5e-10, (st, rel)
    assert st["refinements"] >= 2, st  # FP32 alone cannot reach 1e-10
    # error against the all-FP64 solve, bounded by the conditioning times the tolerance
    np.testing.assert_allclose(x32, x64, rtol=0, atol=1e-7 * np.abs(x64).max())


@pytest.mark.parametrize("precond", ["none", "jacobi", "block jacobi"])
def test_preconditioner_variants(pairs, precond):
    from mrhyde_amd import krylov_opts

    g, o = pairs["2d_q2_nonlinear_perturbed"]
    u = random_state(o, 5)
    ro, jo = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    x, st = g.linearSolver(jo, ro, opts=krylov_opts(solver="gmres", tol=1e-11, max_iters=5000, precond=precond))
    rel = np.linalg.norm(ro - spmv_np(rp, ci, jo, x)) / np.linalg.norm(ro)
    assert st["rel_res"] <= 1e-11 and rel <= 2e-11, (precond, st, rel)


def test_unknown_preconditioner_is_rejected(pairs):
    from mrhyde_amd import MhxError, krylov_opts

    g, o = pairs["2d_q1_poisson_c1shape"]
    ro, jo = o.assembleJacRes(o.setDirichlet(np.zeros(o.n_dof)))
    opts = krylov_opts()
    opts.precond = 7
    with pytest.raises(MhxError):
        g.linearSolver(jo, ro, opts=opts)


def nopts(max_iters, nl_tol, solver, tol=1e-12, backtracking=1):
    o = NewtonOpts()
    o.max_iters, o.nl_tol, o.backtracking = max_iters, nl_tol, backtracking
    o.linear = KrylovOpts()
    o.linear.solver, o.linear.tol, o.linear.max_iters, o.linear.restart = solver, tol, 20000, 50
    o.linear.precond = MHX_PREC_JACOBI
    return o


def test_newton_parity_nonlinear(pairs):
    g, o = pairs["3d_q2_c2shape"]
    u0 = o.setDirichlet(np.zeros(o.n_dof))
    ug, sg = g.nonlinearSolver(u0.copy(), linear=nopts(5, 1e-10, MHX_KRYLOV_GMRES).linear, max_iters=5,
                               nl_tol=1e-10, backtracking=False)
    uo, so = o.nonlinearSolver(u0.copy(), nopts(5, 1e-10, MHX_KRYLOV_GMRES, backtracking=0))
    assert sg["nl_iters"] == so["nl_iters"]
    np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-9 * np.abs(uo).max())
    assert np.abs(o.assembleRes(ug)).max() <= 1e-9 * sg["res_norm0"]


def test_newton_absolute_tolerance_matches_oracle():
    """The absolute nonlinear tolerance on the GPU stops at the oracle's
    iteration, with the oracle's iterate."""
    from mrhyde_amd import Mhx
    from test_oracle import _newton_trace
    from test_oracle import nopts as trace_opts

    o, u0, norms = _newton_trace()
    g = Mhx(o.problem)
    tol = np.sqrt(norms[1] * norms[2])
    opts = trace_opts(max_iters=10, nl_tol=1e-30, backtracking=0, tol=1e-14)
    opts.relative_off, opts.use_abs_tol, opts.abs_tol = 1, 1, tol
    uo, so = o.nonlinearSolver(u0.copy(), opts)
    ug, sg = g.nonlinearSolver(u0.copy(), max_iters=10, nl_tol=1e-30, backtracking=False, use_relative_tol=False,
                               use_absolute_tol=True, abs_tol=tol, linear=opts.linear)
    assert sg["nl_iters"] == so["nl_iters"] == 3
    np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-10 * np.abs(uo).max())
    g.close()


GOLD = json.load(open(os.path.join(GOLDEN, "thermal_gold.json")))


@pytest.mark.parametrize("case", [k for k in GOLD if not k.startswith("_")])
def test_gold_l2_on_gpu(case):
    from mrhyde_amd import Mhx

    from conftest import thermal_gold_problem

    p, true, gd = thermal_gold_problem(GOLD[case])
    g = Mhx(p)
    u = g.setDirichlet(np.zeros(g.n_dof))
    u, st = g.nonlinearSolver(u, max_iters=gd["nl_iters"], nl_tol=gd["nl_tol"], solver="cg", tol=1e-12,
                              linear_max_iters=20000)
    l2, l2g = g.computeError(u, true)
    assert float(f"{l2:.6g}") == gd["gold"]["T"]
    if "grad(T)" in gd["gold"]:
        assert float(f"{l2g:.6g}") == gd["gold"]["grad(T)"]


NS_GOLD = json.load(open(os.path.join(GOLDEN, "ns_gold.json")))
CHANNEL_EXACT = [lambda x: 0.5 * x[:, 1] * (1 - x[:, 1])] + [lambda x: 0.0 * x[:, 0]] * 3


@pytest.mark.parametrize("extrude", [0, 2])
def test_ns_channel_gold_on_gpu(extrude):
    """regression/navierstokes/channel solved on the GPU (Newton, GMRES with
    nodal block Jacobi): L2 errors of ux, uy, pr equal the gold values (2-D),
    or the gold values times sqrt(depth) for the extruded 3-D channel."""
    from conftest import field_l2_errors, ns_channel_problem
    from mrhyde_amd import Mhx
    from oracle import Oracle

    deck = NS_GOLD["navierstokes/channel"]
    p = ns_channel_problem(deck, nz=extrude, depth=0.1 * extrude)
    g = Mhx(p)
    print(f"[ns] channel extrude={extrude} n_dof={g.n_dof}", flush=True)
    u = g.setDirichlet(np.zeros(g.n_dof))
    u, st = g.nonlinearSolver(u, max_iters=10, nl_tol=1e-10, solver="gmres", tol=1e-13, precond="block jacobi",
                              linear_max_iters=20000)
    e = field_l2_errors(Oracle(p), u, CHANNEL_EXACT)
    s = np.sqrt(0.1 * extrude) if extrude else 1.0
    fields = [(0, "ux"), (1, "uy"), (p.dim, "pr")]
    for f, name in fields:
        ref = s * deck["gold"][name]
        assert abs(e[f] - ref) <= 5e-6 * ref, (name, e[f], ref)


def test_ns_newton_parity(pairs):
    """3-D lid-driven cavity (SUPG + PSPG, perturbed mesh): GPU Newton with
    GMRES + block Jacobi against the oracle's Newton."""
    g, o = pairs["3d_ns_lid_cavity"]
    u0 = o.setDirichlet(np.zeros(o.n_dof))
    ug, sg = g.nonlinearSolver(u0.copy(), max_iters=8, nl_tol=1e-10, solver="gmres", tol=1e-12,
                               precond="block jacobi", linear_max_iters=20000)
    uo, so = o.nonlinearSolver(u0.copy(), nopts(8, 1e-10, MHX_KRYLOV_GMRES))
    np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-9 * np.abs(uo).max())
    assert np.abs(o.assembleRes(ug)).max() <= 1e-9 * sg["res_norm0"]


def test_ns_mixed_precision_newton():
    """The mixed-precision solve path on a 3-D cavity: FP32 Jacobian, block-
    Jacobi preconditioner and inner GMRES inside FP64 refinement to 1e-10
    relative residual at every Newton step; the converged state matches the
    all-FP64 Newton (error stated as the bound below)."""
    from mrhyde_amd import Mhx, rng

    p = Problem(dim=3, order=1, quadrature=2, n=(8, 8, 8), physics=MHX_PHYS_NAVIER_STOKES, mu=0.01, supg=True,
                pspg=True, dbc_sides=1 | 2 | 4 | 16 | 32, lid_sides=8, lid_u=(1.0, 0.0, 0.0), pin_pressure=True)
    g = Mhx(p)
    u0 = g.setDirichlet(1e-3 * rng(2026, g.n_dof, 2, 0.0, 1.0))
    kw = dict(max_iters=12, nl_tol=1e-9, solver="gmres", tol=1e-10, linear_max_iters=20000)
    u32, s32 = g.nonlinearSolver(u0.copy(), precond="block jacobi-fp32", **kw)
    u64, s64 = g.nonlinearSolver(u0.copy(), precond="block jacobi", **kw)
    print(f"[ns] mixed: nl {s32['nl_iters']} refinements {s32['refinements']} vs fp64 nl {s64['nl_iters']}", flush=True)
    assert s32["refinements"] >= 2 * s32["nl_iters"] - 2
    assert s32["res_norm"] <= 1e-9 * s32["res_norm0"] and s64["res_norm"] <= 1e-9 * s64["res_norm0"]
    np.testing.assert_allclose(u32, u64, rtol=0, atol=1e-8 * np.abs(u64).max())


# --- failure handling, forced from the host (every kernel still exits normally) ----
def test_nan_input_is_reported(pairs):
    from mrhyde_amd import MhxError

    g, o = pairs["2d_q1_poisson_c1shape"]
    u = o.setDirichlet(np.zeros(o.n_dof))
    u[o.n_dof // 2] = np.nan
    with pytest.raises(MhxError) as ei:
        g.nonlinearSolver(u, max_iters=3, solver="cg")
    assert ei.value.code == -5


def test_singular_jacobian_is_reported():
    from mrhyde_amd import Mhx, MhxError

    p = Problem(dim=2, order=1, quadrature=2, n=(4, 4), k0=np.zeros(16), src_elem=np.ones(16))
    g = Mhx(p)
    u = g.setDirichlet(np.zeros(g.n_dof))
    with pytest.raises(MhxError) as ei:
        g.nonlinearSolver(u, max_iters=2, solver="cg", tol=1e-10)
    assert ei.value.code == -5


def test_iteration_cap_returns(pairs):
    g, o = pairs["3d_q2_c2shape"]
    u = o.setDirichlet(np.zeros(o.n_dof))
    _, st = g.nonlinearSolver(u, max_iters=1, nl_tol=1e-30, solver="gmres", tol=1e-10, linear_max_iters=2000)
    assert st["nl_iters"] == 1 and st["status"] == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_ns_traction_sides_match_oracle(dim):
    """Navier-Stokes with traction on two sides of a perturbed mesh: residual
    and Jacobian on the GPU against the oracle, both schedules."""
    from mrhyde_amd import Mhx
    from oracle import Oracle
    from test_ns_pinning import _traction_problem

    flux = [([0.7, -0.4, 0.25], [0.0, 1.1, 0.7], [np.pi / 2, 0.3, 0.1]), ([0.2, 0.5, -0.3], [0.9, 0.0, 0.0],
                                                                          [0.4, np.pi / 2, np.pi / 2])]
    p = _traction_problem(dim, flux, 2 | 8, perturb=0.15)
    g, o = Mhx(p), Oracle(p)
    u = o.setDirichlet(np.random.default_rng(4).uniform(-0.5, 0.5, o.n_dof))
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    for colouring in (2, 0):
        print(f"[traction] dim={dim} colouring={colouring}", flush=True)
        g.set_option("assembly.colouring", colouring)
        rg, jg = g.assembleJacRes(u)
        assert_entries_close(jg, jo, rp, what=f"traction {dim}-D Jacobian")
        assert_entries_close(rg, ro, None, what=f"traction {dim}-D residual")
    g.close()
