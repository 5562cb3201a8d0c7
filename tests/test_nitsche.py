"""Weak (Nitsche) Dirichlet sides of the thermal module (thermal.cpp:226-281).

The reference holds no gold output for weak Dirichlet data, so the oracle is
pinned by what the mathematics fixes:
  * consistency: a linear solution of the Laplace equation is reproduced to
    round-off on perturbed Q1 and Q2 meshes, 2-D and 3-D, for the symmetric
    and the non-symmetric form;
  * symmetry: with form_param = 1 and a linear k the Jacobian is symmetric;
    with form_param = -1 its skew part is twice the consistency term (so a
    sign error in either term shows);
  * the Jacobian equals central differences of the residual (nonlinear k,
    perturbed mesh, weak, strong and Neumann sides mixed);
  * the nodal error of a manufactured solution with non-zero boundary data
    falls as h^2 (Q1).
GPU: residual and Jacobian within 1e-12 of the oracle, and a Newton solve.
"""
import numpy as np
import pytest

from conftest import assert_entries_close, spmv_np, to_scipy
from mrhyde_amd.abi import MHX_KRYLOV_GMRES, MHX_PREC_JACOBI, NewtonOpts, Problem

PI = np.pi


def all_sides(dim):
    return (1 << (2 * dim)) - 1


def newton_opts():
    o = NewtonOpts()
    o.max_iters, o.nl_tol, o.backtracking = 10, 1e-13, 1
    o.linear.solver, o.linear.precond = MHX_KRYLOV_GMRES, MHX_PREC_JACOBI
    o.linear.max_iters, o.linear.tol, o.linear.restart = 20000, 1e-13, 100
    return o


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_oracle_nitsche_reproduces_linear_solution(dim, order, form):
    from oracle import Oracle

    n = (5, 4, 3)[:dim]
    lin = (2.0, -0.5, 0.3)[:dim]
    p = Problem(dim=dim, order=order, quadrature=2 * order, n=n, dbc_sides=0, wdbc_sides=all_sides(dim),
                wdbc_form=form, wd_const=1.0, wd_lin=lin, perturb=0.2, perturb_seed=3)
    o = Oracle(p)
    u, _ = o.nonlinearSolver(np.zeros(o.n_dof), newton_opts())
    exact = 1.0 + o.dof_coords()[:, :dim] @ np.array(lin)
    assert np.abs(u - exact).max() <= 1e-11


def test_oracle_nitsche_symmetry():
    from oracle import Oracle

    base = dict(dim=2, order=2, quadrature=4, n=(4, 3), dbc_sides=0, wdbc_sides=all_sides(2), perturb=0.2,
                perturb_seed=5, k0=np.exp(0.3 * np.random.default_rng(2).standard_normal(12)))
    u = np.random.default_rng(3).standard_normal(Oracle(Problem(**base)).n_dof)
    mats = {}
    for form in (0, 1, 2):
        o = Oracle(Problem(wdbc_form=form, **base))
        _, j = o.assembleJacRes(u)
        rp, ci = o.crs_pattern()
        mats[form] = to_scipy(rp, ci, j)
    A0, A1, A2 = mats[0], mats[1], mats[2]
    scale = abs(A0).max()
    assert abs(A0 - A0.T).max() <= 1e-14 * scale
    # form_param enters linearly: A(s) = A(0) + s B with B the symmetry term
    assert abs((A0 - A2) + (A1 - A2)).max() <= 1e-14 * scale
    # the consistency term is C = B^T, so A(-1) - A(-1)^T = 2 (C - B) = 2 (B^T - B), not zero
    B = A0 - A2
    assert abs(A1 - A1.T).max() > 1e-3 * scale
    assert abs((A1 - A1.T) - 2 * (B.T - B)).max() <= 1e-13 * scale


@pytest.mark.parametrize("case", ["2d_q2", "3d_q1"])
def test_oracle_nitsche_jacobian_vs_fd(case):
    from oracle import Oracle

    if case == "2d_q2":
        p = Problem(dim=2, order=2, quadrature=4, n=(4, 3), diff_u2=1.0, dbc_sides=1, wdbc_sides=2 | 4,
                    neumann_sides=8, neumann_flux=[((0.0, 1.0), (PI, PI), (0.0, 0.0))], wd_const=0.3,
                    wd_lin=(0.5, -0.2), wd_modes=[(0.4, (PI, 2 * PI), (0.1, 0.2))], perturb=0.25, perturb_seed=4,
                    k0=np.exp(0.3 * np.random.default_rng(6).standard_normal(12)))
    else:
        p = Problem(dim=3, order=1, quadrature=2, n=(3, 3, 2), diff_u2=0.7, dbc_sides=16, wdbc_sides=1 | 2 | 8 | 32,
                    wd_const=-0.2, wd_lin=(0.1, 0.2, 0.3), perturb=0.2, perturb_seed=7, wdbc_form=1)
    o = Oracle(p)
    rng = np.random.default_rng(8)
    u = o.setDirichlet(0.5 * rng.standard_normal(o.n_dof))
    _, j = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    v = rng.standard_normal(o.n_dof)
    mask, _ = o.dirichlet()
    v[mask.astype(bool)] = 0.0
    eps = 1e-6
    # residual = -F, Jacobian = +dF/du
    fd = -(o.assembleRes(u + eps * v) - o.assembleRes(u - eps * v)) / (2 * eps)
    jv = spmv_np(rp, ci, j, v)
    free = ~mask.astype(bool)
    np.testing.assert_allclose(jv[free], fd[free], rtol=0, atol=1e-6 * np.abs(jv).max())


def test_oracle_nitsche_convergence_rate():
    """T = sin(pi x) sin(pi y) + x, weak data g = x on every side: the nodal
    error falls as h^2."""
    from oracle import Oracle

    errs = []
    for n in (8, 16, 32):
        p = Problem(dim=2, order=1, quadrature=2, n=(n, n), dbc_sides=0, wdbc_sides=all_sides(2),
                    wd_lin=(1.0, 0.0), modes=[(2 * PI * PI, (PI, PI))])
        o = Oracle(p)
        u, _ = o.nonlinearSolver(np.zeros(o.n_dof), newton_opts())
        X = o.dof_coords()
        errs.append(np.abs(u - (np.sin(PI * X[:, 0]) * np.sin(PI * X[:, 1]) + X[:, 0])).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.8), (errs, rates)


# ---- GPU ---------------------------------------------------------------------
GPU_CASES = {
    "2d_q1_all_weak": Problem(dim=2, order=1, quadrature=2, n=(9, 7), dbc_sides=0, wdbc_sides=15, wd_const=0.5,
                              wd_lin=(1.0, -2.0), modes=[(2 * PI * PI, (PI, PI))]),
    "2d_q2_mixed_nonlinear": Problem(dim=2, order=2, quadrature=4, n=(6, 5), diff_u2=1.0, dbc_sides=1,
                                     wdbc_sides=2 | 4, neumann_sides=8,
                                     neumann_flux=[((0.0, 1.0), (PI, PI), (0.0, 0.0))], wd_const=0.3,
                                     wd_modes=[(0.4, (PI, 2 * PI), (0.1, 0.2))], perturb=0.25, perturb_seed=4,
                                     k0=np.exp(0.3 * np.random.default_rng(6).standard_normal(30))),
    "2d_q3_nonsymmetric": Problem(dim=2, order=3, quadrature=6, n=(4, 3), dbc_sides=0, wdbc_sides=15, wdbc_form=1,
                                  wd_lin=(0.3, 0.7), perturb=0.2, perturb_seed=9),
    "2d_q2_incomplete": Problem(dim=2, order=2, quadrature=4, n=(5, 4), dbc_sides=4 | 8, wdbc_sides=1 | 2,
                                wdbc_form=2, wd_const=1.0, diff_u2=0.5),
    "3d_q1_perturbed": Problem(dim=3, order=1, quadrature=2, n=(4, 3, 3), diff_u2=0.7, dbc_sides=16,
                               wdbc_sides=1 | 2 | 8 | 32, wd_const=-0.2, wd_lin=(0.1, 0.2, 0.3), perturb=0.2,
                               perturb_seed=7, wdbc_form=1),
    "3d_q2_all_weak": Problem(dim=3, order=2, quadrature=4, n=(3, 3, 2), dbc_sides=0, wdbc_sides=63, diff_u2=1.0,
                              wd_modes=[(0.5, (PI, PI, PI), (0.3, 0.0, 0.1))],
                              src_elem=np.random.default_rng(3).uniform(0, 1, 18)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_nitsche_assembly_parity(name):
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = GPU_CASES[name]
    g, o = Mhx(p), Oracle(p)
    u = o.setDirichlet(np.random.default_rng(11).standard_normal(o.n_dof))
    print(f"[nitsche] {name} n_dof={o.n_dof}", flush=True)
    rg, jg = np.zeros(g.n_dof), np.zeros(g.nnz)
    g.assemble(u, rg, jg)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"{name} Jacobian")
    assert_entries_close(rg, ro, None, what=f"{name} residual")
    assert_entries_close(g.assembleRes(u), ro, None, what=f"{name} residual-only")


@pytest.mark.gpu
def test_nitsche_newton_solve():
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = GPU_CASES["2d_q2_mixed_nonlinear"]
    g, o = Mhx(p), Oracle(p)
    u0 = o.setDirichlet(np.zeros(o.n_dof))
    ug, st = g.nonlinearSolver(u0.copy(), max_iters=10, nl_tol=1e-12, solver="gmres", tol=1e-12,
                               linear_max_iters=5000)
    uo, _ = o.nonlinearSolver(u0.copy(), newton_opts())
    np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-9 * max(1.0, np.abs(uo).max()))
