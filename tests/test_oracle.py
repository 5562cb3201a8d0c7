"""Pins the CPU oracle against the reference's own known answers, closed
forms and finite differences (no GPU)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import GOLDEN, spmv_np, to_scipy
from mrhyde_amd.abi import (MHX_KRYLOV_CG, MHX_KRYLOV_GMRES, KrylovOpts, NewtonOpts, Problem)
from oracle import Oracle, gauss

PI = np.pi
# Intrepid2 vertex order expressed in the oracle's lexicographic local order
PERM_QUAD = [0, 1, 3, 2]
PERM_HEX = [0, 1, 3, 2, 4, 5, 7, 6]


def kopts(solver=MHX_KRYLOV_CG, tol=1e-12, max_iters=20000, restart=50):
    o = KrylovOpts()
    o.solver, o.tol, o.max_iters, o.restart = solver, tol, max_iters, restart
    return o


def nopts(max_iters=2, nl_tol=1e-7, backtracking=1, **kw):
    o = NewtonOpts()
    o.max_iters, o.nl_tol, o.backtracking = max_iters, nl_tol, backtracking
    o.linear = kopts(**kw)
    return o


# --- reference known-answer test: regression/discretization/HGRAD ---------------
@pytest.mark.parametrize("topo", ["quad", "hex"])
def test_hgrad_kat(topo):
    kat = json.load(open(os.path.join(GOLDEN, "hgrad_kat.json")))[topo]
    dim = 2 if topo == "quad" else 3
    perm = PERM_QUAD if dim == 2 else PERM_HEX
    o = Oracle(Problem(dim=dim, order=1, quadrature=2, n=(1, 1, 1), dbc_sides=0))
    phi, grad, wts, _ = o.elem_basis(0)
    gphi = np.array(kat["phi"])
    ggrad = np.array(kat["grad"])
    for d in range(len(perm)):
        np.testing.assert_allclose(phi[perm[d]], gphi[d], rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(grad[perm[d]], ggrad[d], rtol=2e-6, atol=1e-7)
    assert wts.sum() == pytest.approx(1.0, rel=1e-14)


# --- reference gold L2 errors: regression/thermal/*/mrhyde.gold ----------------------
GOLD = json.load(open(os.path.join(GOLDEN, "thermal_gold.json")))
CASES = [k for k in GOLD if not k.startswith("_")]


def gold_problem(case):
    from conftest import thermal_gold_problem

    return thermal_gold_problem(GOLD[case])


@pytest.mark.parametrize("case", CASES)
def test_thermal_gold_l2(case):
    p, true, g = gold_problem(case)
    o = Oracle(p)
    u = o.setDirichlet(np.zeros(o.n_dof))
    u, st = o.nonlinearSolver(u, nopts(max_iters=g["nl_iters"], nl_tol=g["nl_tol"]))
    l2, l2g = o.computeError(u, true)
    assert float(f"{l2:.6g}") == g["gold"]["T"], (l2, g["gold"]["T"])
    if "grad(T)" in g["gold"]:
        assert float(f"{l2g:.6g}") == g["gold"]["grad(T)"], (l2g, g["gold"]["grad(T)"])


# --- closed forms ------------------------------------------------------------------
def test_q1_quad_element_stiffness():
    # unit square, k = 1: the standard bilinear stiffness matrix (lexicographic order)
    o = Oracle(Problem(dim=2, order=1, quadrature=2, n=(1, 1), dbc_sides=0))
    _, jac = o.assembleJacRes(np.zeros(4))
    rp, ci = o.crs_pattern()
    K = to_scipy(rp, ci, jac).toarray()
    a, b, c = 2 / 3, -1 / 6, -1 / 3
    ref = np.array([[a, b, b, c], [b, a, c, b], [b, c, a, b], [c, b, b, a]])
    np.testing.assert_allclose(K, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dim,p,n", [(2, 1, 5), (2, 2, 4), (2, 3, 3), (3, 1, 4), (3, 2, 3), (2, 4, 2)])
def test_crs_nnz_closed_form(dim, p, n):
    o = Oracle(Problem(dim=dim, order=p, quadrature=2 * p, n=(n,) * dim))
    one_d = (n - 1) * (2 * p + 1) + 2 * (p + 1) + n * (p - 1) * (p + 1)
    assert o.nnz == one_d ** dim
    rp, ci = o.crs_pattern()
    for r in range(o.n_dof):
        row = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(row) > 0) and r in row


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_gauss_rule_exactness(n):
    x, w = gauss(n)
    assert np.all(np.diff(x) < 0)  # descending, the order of the HGRAD KAT
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert np.dot(w, x ** k) == pytest.approx(exact, abs=1e-14)


@pytest.mark.parametrize("dim,p,q", [(2, 2, 4), (3, 2, 4), (2, 4, 8)])
def test_basis_invariants_perturbed(dim, p, q):
    o = Oracle(Problem(dim=dim, order=p, quadrature=q, n=(3,) * dim, perturb=0.25, perturb_seed=3))
    for e in range(o.n_elem):
        phi, grad, wts, _ = o.elem_basis(e)
        np.testing.assert_allclose(phi.sum(axis=0), 1.0, atol=1e-13)
        np.testing.assert_allclose(grad.sum(axis=0), 0.0, atol=1e-11)
    # total volume of the box is preserved by the interior jitter
    vol = sum(o.elem_basis(e)[2].sum() for e in range(o.n_elem))
    assert vol == pytest.approx(1.0, rel=1e-13)


def test_ad_size_selection():
    # assemblyManager_construct.hpp:128-171
    sizes = {(2, 1): 4, (3, 1): 8, (2, 2): 16, (3, 2): 32, (2, 4): 32, (2, 3): 16}
    for (dim, p), ad in sizes.items():
        assert Oracle(Problem(dim=dim, order=p, quadrature=2 * p, n=(1,) * dim)).ad_size == ad


# --- dual-number Jacobian vs central differences -----------------------------------------
@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_jacobian_vs_fd(dim, p):
    ne = 2 ** dim
    rng = np.random.default_rng(11)
    prob = Problem(dim=dim, order=p, quadrature=2 * p, n=(2,) * dim, diff_u2=1.0,
                   k0=np.exp(0.5 * rng.standard_normal(ne)), src_elem=rng.uniform(0, 1, ne),
                   modes=[(1.3, [PI, 2 * PI, PI])], perturb=0.2, perturb_seed=5, dbc_sides=1 | 2)
    o = Oracle(prob)
    u = rng.standard_normal(o.n_dof)
    res, jac = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    J = to_scipy(rp, ci, jac).toarray()
    mask, _ = o.dirichlet()
    h = 1e-6
    for j in range(o.n_dof):
        e = np.zeros(o.n_dof)
        e[j] = h
        col = -(o.assembleRes(u + e) - o.assembleRes(u - e)) / (2 * h)  # res = -F
        free = mask == 0
        np.testing.assert_allclose(J[free, j], col[free], rtol=1e-6, atol=1e-7 * np.abs(J).max())
    # Dirichlet rows are identity rows, residual zero there
    fixed = np.where(mask == 1)[0]
    np.testing.assert_array_equal(J[fixed][:, fixed], np.eye(len(fixed)))
    assert np.all(res[fixed] == 0.0)
    assert np.abs(J[fixed]).sum() == len(fixed)


def test_linear_operator_properties():
    # linear k: interior rows of the diffusion operator sum to zero and are symmetric
    rng = np.random.default_rng(2)
    prob = Problem(dim=3, order=2, quadrature=4, n=(2, 2, 2), k0=np.exp(rng.standard_normal(8)),
                   perturb=0.2, perturb_seed=1, dbc_sides=0)
    o = Oracle(prob)
    res0, jac = o.assembleJacRes(np.zeros(o.n_dof))
    rp, ci = o.crs_pattern()
    J = to_scipy(rp, ci, jac).toarray()
    np.testing.assert_allclose(J.sum(axis=1), 0.0, atol=1e-13)
    np.testing.assert_allclose(J, J.T, atol=1e-14)
    u = rng.standard_normal(o.n_dof)
    # affine residual: res(u) = res(0) - J u
    np.testing.assert_allclose(o.assembleRes(u), res0 - J @ u, atol=1e-12)


# --- Krylov against a direct solve --------------------------------------------------------
@pytest.mark.parametrize("solver", [MHX_KRYLOV_CG, MHX_KRYLOV_GMRES])
def test_krylov_vs_spsolve(solver):
    rng = np.random.default_rng(4)
    prob = Problem(dim=2, order=2, quadrature=4, n=(6, 5), k0=np.exp(rng.standard_normal(30)),
                   src_elem=rng.uniform(0, 1, 30))
    o = Oracle(prob)
    res, jac = o.assembleJacRes(o.setDirichlet(np.zeros(o.n_dof)))
    rp, ci = o.crs_pattern()
    A = to_scipy(rp, ci, jac)
    x_direct = spla.spsolve(A.tocsc(), res)
    x, st = o.linearSolver(jac, res, opts=kopts(solver=solver, tol=1e-11))
    assert st["rel_res"] <= 1e-11
    r = res - spmv_np(rp, ci, jac, x)
    assert np.linalg.norm(r) / np.linalg.norm(res) <= 1e-11
    np.testing.assert_allclose(x, x_direct, rtol=0, atol=1e-9 * np.abs(x_direct).max())


def test_newton_nonlinear_quadratic_convergence():
    rng = np.random.default_rng(8)
    prob = Problem(dim=2, order=2, quadrature=4, n=(4, 4), diff_u2=1.0, k0=np.exp(0.5 * rng.standard_normal(16)),
                   src_elem=rng.uniform(0, 20, 16), dbc_sides=1 | 2)
    o = Oracle(prob)
    norms = []
    for k in range(1, 7):
        u = o.setDirichlet(np.zeros(o.n_dof))
        u, st = o.nonlinearSolver(u, nopts(max_iters=k, nl_tol=1e-30, backtracking=0, tol=1e-14))
        norms.append(np.abs(o.assembleRes(u)).max())
    # residual falls super-linearly once in the basin
    assert norms[-1] < 1e-10 * norms[0]
    assert norms[3] < norms[2] ** 1.5 / max(norms[0], 1) * 10 or norms[3] < 1e-12


def _newton_trace():
    """The nonlinear problem, and ||F(u_j)||_inf after j = 0..6 Newton updates."""
    rng = np.random.default_rng(8)
    prob = Problem(dim=2, order=2, quadrature=4, n=(4, 4), diff_u2=1.0, k0=np.exp(0.5 * rng.standard_normal(16)),
                   src_elem=rng.uniform(0, 20, 16), dbc_sides=1 | 2)
    o = Oracle(prob)
    u0 = o.setDirichlet(np.zeros(o.n_dof))
    norms = [np.abs(o.assembleRes(u0)).max()]
    for k in range(1, 7):
        u, _ = o.nonlinearSolver(u0.copy(), nopts(max_iters=k, nl_tol=1e-30, backtracking=0, tol=1e-14))
        norms.append(np.abs(o.assembleRes(u)).max())
    return o, u0, norms


def test_newton_absolute_tolerance():
    """"use relative TOL" false with "use absolute TOL" (solverManager_
    solvers.hpp:430-445): Newton stops at the first iterate whose residual
    inf-norm is below the absolute tolerance, found here from the
    per-iteration norms of separate runs; without either test it runs
    "max nonlinear iters" steps; the default absolute tolerance is
    min(nonlinear TOL, 1e-6)."""
    o, u0, norms = _newton_trace()
    assert all(norms[j + 1] < norms[j] for j in range(1, 5))  # (the first full step overshoots)
    for j in (2, 3, 4):
        tol = np.sqrt(norms[j - 1] * norms[j])  # between iterates j-1 and j, below every earlier one
        assert min(norms[:j]) > tol
        opts = nopts(max_iters=10, nl_tol=1e-30, backtracking=0, tol=1e-14)
        opts.relative_off, opts.use_abs_tol, opts.abs_tol = 1, 1, tol
        u, st = o.nonlinearSolver(u0.copy(), opts)
        assert st["nl_iters"] == j + 1
        assert st["res_norm"] == pytest.approx(norms[j], rel=1e-12)
    # neither test: exactly max_iters steps although the residual is tiny
    opts = nopts(max_iters=6, nl_tol=1.0, backtracking=0, tol=1e-14)
    opts.relative_off = 1
    _, st = o.nonlinearSolver(u0.copy(), opts)
    assert st["nl_iters"] == 6
    # default absolute tolerance min(nl_tol, 1e-6), not nl_tol
    opts = nopts(max_iters=10, nl_tol=1.0, backtracking=0, tol=1e-14)
    opts.relative_off, opts.use_abs_tol = 1, 1
    _, st = o.nonlinearSolver(u0.copy(), opts)
    j = next(k for k, v in enumerate(norms) if v < 1e-6)
    assert st["nl_iters"] == j + 1


# --- navier stokes ----------------------------------------------------------------------
from conftest import field_l2_errors, ns_channel_problem  # noqa: E402
from mrhyde_amd.abi import MHX_PHYS_NAVIER_STOKES  # noqa: E402

NS_GOLD = json.load(open(os.path.join(GOLDEN, "ns_gold.json")))
CHANNEL_EXACT = [lambda x: 0.5 * x[:, 1] * (1 - x[:, 1])] + [lambda x: 0.0 * x[:, 0]] * 3


def test_ns_channel_gold():
    """regression/navierstokes/channel: 2-D Q1-Q1 PSPG channel, L2 errors of
    ux, uy and pr reproduced to the printed digits."""
    deck = NS_GOLD["navierstokes/channel"]
    o = Oracle(ns_channel_problem(deck))
    u = o.setDirichlet(np.zeros(o.n_dof))
    u, st = o.nonlinearSolver(u, nopts(max_iters=10, nl_tol=1e-10, solver=MHX_KRYLOV_GMRES, tol=1e-13))
    e = field_l2_errors(o, u, CHANNEL_EXACT)
    gold = deck["gold"]
    for f, name in enumerate(["ux", "uy", "pr"]):
        assert abs(e[f] - gold[name]) <= 5e-6 * gold[name], (name, e[f], gold[name])


def test_ns_channel_extruded_3d():
    """The same channel extruded in z (2 cubes of 0.1 deep, natural conditions on
    the z faces, uz = 0 there): the 3-D discrete solution is the 2-D one, so the L2
    errors are the 2-D gold values times sqrt(depth) (element size, and so tau,
    is the same 0.1 in both)."""
    deck = NS_GOLD["navierstokes/channel"]
    o = Oracle(ns_channel_problem(deck, nz=2, depth=0.2))
    u = o.setDirichlet(np.zeros(o.n_dof))
    u, st = o.nonlinearSolver(u, nopts(max_iters=10, nl_tol=1e-10, solver=MHX_KRYLOV_GMRES, tol=1e-13))
    e = field_l2_errors(o, u, CHANNEL_EXACT)
    gold = deck["gold"]
    s = np.sqrt(0.2)
    for f, name in [(0, "ux"), (1, "uy"), (3, "pr")]:
        assert abs(e[f] - s * gold[name]) <= 5e-6 * s * gold[name], (name, e[f], s * gold[name])
    assert e[2] < 1e-10  # uz


@pytest.mark.parametrize("dim", [2, 3])
def test_ns_jacobian_vs_fd(dim):
    """SUPG + PSPG element Jacobian (32-slot duals in 3-D) against central
    differences of the residual on a perturbed mesh with lid, no-slip and a
    pinned pressure."""
    rng = np.random.default_rng(31 + dim)
    prob = Problem(dim=dim, order=1, quadrature=2, n=(2,) * dim, physics=MHX_PHYS_NAVIER_STOKES,
                   rho=1.3, mu=0.05, supg=True, pspg=True, force=(0.4, -0.2, 0.3), dbc_sides=1 | 4,
                   lid_sides=8, lid_u=(1.0, 0.0, 0.0), pin_pressure=True, perturb=0.2, perturb_seed=3)
    o = Oracle(prob)
    u = o.setDirichlet(rng.standard_normal(o.n_dof))
    res, jac = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    J = to_scipy(rp, ci, jac).toarray()
    mask, vals = o.dirichlet()
    h = 1e-6
    free = mask == 0
    for j in range(o.n_dof):
        e = np.zeros(o.n_dof)
        e[j] = h
        col = -(o.assembleRes(u + e) - o.assembleRes(u - e)) / (2 * h)
        np.testing.assert_allclose(J[free, j], col[free], rtol=1e-6, atol=1e-7 * np.abs(J).max())
    fixed = np.where(mask == 1)[0]
    np.testing.assert_array_equal(J[fixed][:, fixed], np.eye(len(fixed)))
    assert np.all(res[fixed] == 0.0)
    # lid and no-slip values; pressure pinned at node 0 (DOF dim)
    assert mask[dim] == 1 and vals[dim] == 0.0
    assert set(np.round(vals[fixed], 12)) <= {0.0, 1.0}


@pytest.mark.parametrize("dim,n", [(2, (4, 3)), (3, (3, 2, 2))])
def test_ns_crs_nnz_closed_form(dim, n):
    o = Oracle(Problem(dim=dim, order=1, quadrature=2, n=n, physics=MHX_PHYS_NAVIER_STOKES))
    nf = dim + 1
    assert o.n_dof == nf * int(np.prod([k + 1 for k in n]))
    assert o.nnz == nf * nf * int(np.prod([3 * k + 1 for k in n]))
