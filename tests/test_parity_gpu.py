"""GPU parity: libmhx.so (HIP, gfx950) against the CPU oracle on the same
seeded inputs.  Bars: CRS pattern bit-exact; residual and Jacobian entries
within 1e-12 relative (plus 1e-14 of the row's largest entry for entries that
cancel to ~0); Krylov solutions at <= 1e-10 true relative residual."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_entries_close, spmv_np, to_scipy
from mrhyde_amd.abi import MHX_KRYLOV_CG, MHX_KRYLOV_GMRES, KrylovOpts, NewtonOpts, Problem

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


@pytest.mark.parametrize("colouring", [1, 0])
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


@pytest.mark.parametrize("colouring", [1, 0])
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


def nopts(max_iters, nl_tol, solver, tol=1e-12, backtracking=1):
    o = NewtonOpts()
    o.max_iters, o.nl_tol, o.backtracking = max_iters, nl_tol, backtracking
    o.linear = KrylovOpts()
    o.linear.solver, o.linear.tol, o.linear.max_iters, o.linear.restart = solver, tol, 20000, 50
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


GOLD = json.load(open(os.path.join(GOLDEN, "thermal_gold.json")))


@pytest.mark.parametrize("case", [k for k in GOLD if not k.startswith("_")])
def test_gold_l2_on_gpu(case):
    from mrhyde_amd import Mhx

    gd = GOLD[case]
    modes = [(a * PI * PI, [w * PI for w in f]) for a, f in gd["source"]]
    true = [(a, [w * PI for w in f]) for a, f in gd["true"]]
    g = Mhx(Problem(dim=gd["dim"], order=gd["order"], quadrature=gd["quadrature"], n=tuple(gd["n"]), modes=modes))
    u = g.setDirichlet(np.zeros(g.n_dof))
    u, st = g.nonlinearSolver(u, max_iters=gd["nl_iters"], nl_tol=gd["nl_tol"], solver="cg", tol=1e-12,
                              linear_max_iters=20000)
    l2, l2g = g.computeError(u, true)
    assert float(f"{l2:.6g}") == gd["gold"]["T"]
    if "grad(T)" in gd["gold"]:
        assert float(f"{l2g:.6g}") == gd["gold"]["grad(T)"]


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
