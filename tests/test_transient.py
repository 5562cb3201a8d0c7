"""Transient thermal (backward differences, BWE tableau).

CPU: the oracle reproduces the reference's 2D_verification_transient gold
(L2 error of every one of the 20 steps), is exact for a solution linear in
time at every BDF order, and converges at order k for BDF-k (errors against
a time-converged run, exact history, so only the weights are tested).
GPU: transient residual and Jacobian within 1e-12 of the oracle; the gold on
the GPU; a multi-step BDF3 run with the reference's startup schedule equal to
the oracle's.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_entries_close
from mrhyde_amd.abi import MHX_KRYLOV_CG, MHX_PREC_JACOBI, NewtonOpts, Problem, bdf_schedule

PI = np.pi
GOLD = json.load(open(os.path.join(GOLDEN, "transient_gold.json")))["2D_verification_transient"]


def gold_problem(t):
    amp = GOLD["source_sin_amp_pi2"] * PI * PI * np.sin(GOLD["omega_pi"] * PI * t) + \
        GOLD["source_cos_amp_pi"] * PI * np.cos(GOLD["omega_pi"] * PI * t)
    w = [f * PI for f in GOLD["space_freq_pi"]]
    return Problem(dim=GOLD["dim"], order=GOLD["order"], quadrature=GOLD["quadrature"], n=tuple(GOLD["n"]),
                   workset_size=GOLD["workset"], modes=[(amp, w)])


def gold_true(t):
    return [(np.sin(GOLD["omega_pi"] * PI * t), [f * PI for f in GOLD["space_freq_pi"]])]


def newton_opts():
    o = NewtonOpts()
    o.max_iters, o.nl_tol, o.backtracking = 10, 1e-12, 1
    o.linear.solver, o.linear.precond, o.linear.max_iters, o.linear.tol = MHX_KRYLOV_CG, MHX_PREC_JACOBI, 10000, 1e-12
    return o


def check_gold(errs):
    gold = [(t, e) for t, e in GOLD["gold"] if t > 0]
    assert len(errs) == len(gold) == GOLD["steps"]
    for (t, e), (tg, eg) in zip(errs, gold):
        assert abs(t - tg) < 1e-9
        # the gold prints 6 significant digits
        assert abs(e - eg) <= 5e-6 * eg + 1e-12, (t, e, eg)


def test_bdf_schedule_follows_startup_rule():
    assert bdf_schedule(1, 4) == [1, 1, 1, 1]
    assert bdf_schedule(3, 6, startup_order=1) == [1, 1, 1, 3, 3, 3]
    assert bdf_schedule(2, 4, startup_order=1, startup_steps=1) == [1, 2, 2, 2]
    assert bdf_schedule(2, 3) == [2, 2, 2]


def test_oracle_transient_gold():
    from oracle import Oracle

    o = Oracle(gold_problem(0.0))
    errs = []
    dt = GOLD["final_time"] / GOLD["steps"]
    o.transientSolver(np.zeros(o.n_dof), dt, GOLD["steps"], order=GOLD["bdf_order"], at_time=gold_problem,
                      record=lambda t, u: errs.append((t, o.computeError(u, gold_true(t))[0])), opts=newton_opts())
    check_gold(errs)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_oracle_exact_for_linear_in_time(order):
    """T = t (source 1, T = t on the boundary, T(0) = 0): every BDF order with
    a BDF1 start is exact, in space and in time."""
    from oracle import Oracle

    def prob(t):
        return Problem(dim=2, order=2, quadrature=4, n=(5, 4), src_elem=np.ones(20), dbc_const=t)

    o = Oracle(prob(0.0))
    worst = [0.0]

    def rec(t, u):
        worst[0] = max(worst[0], np.abs(u - t).max())

    o.transientSolver(np.zeros(o.n_dof), 0.1, 8, order=order, startup_order=1, at_time=prob, record=rec,
                      opts=newton_opts())
    assert worst[0] <= 1e-11, worst[0]


def _run_exact_history(o, prob, exact_nodal, order, dt, nsteps):
    hist = [exact_nodal(-k * dt) for k in range(order)]
    u = hist[0].copy()
    t = 0.0
    for _ in range(nsteps):
        t += dt
        o.update(prob(t))
        o.set_time_state(dt, order, hist)
        o.setDirichlet(u)
        u, _ = o.nonlinearSolver(u, newton_opts())
        hist = [u.copy()] + hist[:-1]
    o.set_time_state(dt, 0, [])
    return u


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_oracle_bdf_convergence_order(order):
    """T = sin(2 pi t), uniform in space, no Dirichlet sides (zero flux), source
    2 pi cos(2 pi t): the discrete system reduces exactly to the ODE
    T' = f(t), so with exact history the error of BDF-k falls by 2^k when dt
    halves; a wrong weight breaks the rate."""
    from oracle import Oracle

    def prob(t):
        return Problem(dim=2, order=2, quadrature=4, n=(3, 2), dbc_sides=0,
                       modes=[(2 * PI * np.cos(2 * PI * t), (0.0, 0.0))])

    o = Oracle(prob(0.0))
    ones = np.ones(o.n_dof)
    T_end = 0.25
    errs = []
    for n in (10, 20, 40):
        u = _run_exact_history(o, prob, lambda t: np.sin(2 * PI * t) * ones, order, T_end / n, n)
        errs.append(np.abs(u - np.sin(2 * PI * T_end)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - order) < 0.3), (order, errs, rates)


# ---- GPU ---------------------------------------------------------------------
TR_CONFIGS = {
    "2d_q2_nonlinear_perturbed": Problem(dim=2, order=2, quadrature=4, n=(7, 5), diff_u2=1.0,
                                         k0=np.exp(0.5 * np.random.default_rng(1).standard_normal(35)),
                                         modes=[(3.0, (PI, 2 * PI))], perturb=0.3, perturb_seed=2, dbc_sides=1 | 2,
                                         density=1.3, specific_heat=0.7),
    "3d_q2": Problem(dim=3, order=2, quadrature=4, n=(4, 3, 3), diff_u2=0.5,
                     src_elem=np.random.default_rng(5).uniform(0, 1, 36), dbc_sides=1 | 2),
    "3d_q1_perturbed": Problem(dim=3, order=1, quadrature=2, n=(5, 4, 3), modes=[(2.0, (PI, PI, PI))],
                               perturb=0.2, perturb_seed=3),
}


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("name", list(TR_CONFIGS))
def test_transient_assembly_parity(name, order):
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = TR_CONFIGS[name]
    g, o = Mhx(p), Oracle(p)
    rng = np.random.default_rng(7 + order)
    hist = [rng.standard_normal(o.n_dof) for _ in range(order)]
    u = o.setDirichlet(rng.standard_normal(o.n_dof))
    print(f"[transient] {name} order={order} n_dof={o.n_dof}", flush=True)
    g.set_option("assembly.colouring", 2)  # the transient kernel runs element colouring whatever the option
    g.set_time_state(0.05, order, hist)
    o.set_time_state(0.05, order, hist)
    rg = np.zeros(g.n_dof)
    jg = np.zeros(g.nnz)
    g.assemble(u, rg, jg)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"{name} transient Jacobian")
    assert_entries_close(rg, ro, None, what=f"{name} transient residual")
    # back to steady: the steady assembly is unchanged
    g.set_time_state(0.05, 0)
    o.set_time_state(0.05, 0, [])
    rg2 = np.zeros(g.n_dof)
    jg2 = np.zeros(g.nnz)
    g.assemble(u, rg2, jg2)
    ro2, jo2 = o.assembleJacRes(u)
    assert_entries_close(jg2, jo2, rp, what=f"{name} steady Jacobian")
    g.set_option("assembly.colouring", 0)


@pytest.mark.gpu
def test_transient_gold_on_gpu():
    from mrhyde_amd import Mhx

    g = Mhx(gold_problem(0.0))
    errs = []
    dt = GOLD["final_time"] / GOLD["steps"]
    g.transientSolver(np.zeros(g.n_dof), dt, GOLD["steps"], order=GOLD["bdf_order"], at_time=gold_problem,
                      record=lambda t, u: errs.append((t, g.computeError(u, gold_true(t))[0])))
    check_gold(errs)


@pytest.mark.gpu
def test_transient_bdf3_matches_oracle():
    """Nonlinear k(T), BDF3 with a BDF1 start, 8 steps: the GPU and the oracle
    trajectories agree to the Newton / Krylov tolerances."""
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = TR_CONFIGS["2d_q2_nonlinear_perturbed"]
    g, o = Mhx(p), Oracle(p)
    traj_g, traj_o = [], []
    u0 = np.zeros(o.n_dof)
    g.transientSolver(u0, 0.05, 8, order=3, startup_order=1, record=lambda t, u: traj_g.append(u.copy()))
    o.transientSolver(u0, 0.05, 8, order=3, startup_order=1, record=lambda t, u: traj_o.append(u.copy()),
                      opts=newton_opts())
    for a, b in zip(traj_g, traj_o):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * max(1.0, np.abs(b).max()))


# --- multi-stage Butcher tableaus (DIRK, CN) ------------------------------------------
# No reference gold uses them outside multiscale decks: the oracle is pinned by
# exactness for polynomials in time on the ODE limit (a spatially uniform
# solution with no boundary data: M T' = M f(t) node by node), where a stage
# method of order p reduces to a quadrature rule, and by the measured order
# of convergence on T' = -T, where the stages couple through A.
TABLEAU_ORDER = {"BWE": 1, "DIRK-1,2": 2, "CN": 2, "DIRK-2,2": 2, "DIRK-2,3": 3, "DIRK-3,3": 3}


def ode_problem(source, t=0.0):
    return Problem(dim=2, order=1, quadrature=2, n=(2, 2), dbc_sides=0, functions={"thermal source": source},
                   time=t)


def ode_opts():
    o = newton_opts()
    o.max_iters, o.nl_tol = 8, 1e-14
    o.linear.tol = 1e-14
    return o


@pytest.mark.parametrize("tableau,deg", [("BWE", 1), ("DIRK-1,2", 2), ("CN", 2), ("DIRK-2,2", 2),
                                         ("DIRK-2,3", 4), ("DIRK-3,3", 3)])
def test_oracle_stage_methods_exact_for_polynomials(tableau, deg):
    from oracle import Oracle

    # T = t^deg, f = deg t^(deg-1): the quadrature of each tableau is exact to
    # its order (DIRK-2,3 is the 2-point Gauss rule: exact for t^4; DIRK-3,3's
    # diagonal is a 10-digit constant, so its conditions hold to ~1e-10)
    src = "1" if deg == 1 else f"{deg}*t^{deg - 1}"
    o = Oracle(ode_problem(src))
    dt, n = 0.1, 5
    u = o.transientSolver(np.zeros(o.n_dof), dt, n, tableau=tableau, opts=ode_opts(),
                          at_time=lambda t: ode_problem(src, t))
    tol = 1e-9 if tableau == "DIRK-3,3" else 1e-12
    np.testing.assert_allclose(u, (n * dt) ** deg, rtol=tol, atol=tol)


@pytest.mark.parametrize("tableau", list(TABLEAU_ORDER))
def test_oracle_stage_methods_convergence_order(tableau):
    from oracle import Oracle

    # T' = -T, T(0) = 1, errors at t = 0.8 against exp(-0.8)
    o = Oracle(ode_problem("-T"))
    errs = []
    for n in (8, 16, 32):
        u = o.transientSolver(np.ones(o.n_dof), 0.8 / n, n, tableau=tableau, opts=ode_opts())
        errs.append(abs(u[0] - np.exp(-0.8)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    p = TABLEAU_ORDER[tableau]
    assert np.all(np.abs(rates - p) < 0.2), (tableau, errs, rates)


def test_butcher_tableau_checks():
    from mrhyde_amd.abi import butcher_tableau

    for name in TABLEAU_ORDER:
        A, b, c = butcher_tableau(name)
        assert np.sum(b) == pytest.approx(1.0, abs=1e-12)   # consistency
        np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-9)  # row sums = c (CN included)
    with pytest.raises(ValueError):
        butcher_tableau("RK-4,4")
    with pytest.raises(ValueError):
        butcher_tableau("custom", [[0.5, 0.5], [0.0, 0.5]], [0.5, 0.5], [1.0, 0.5])  # not lower triangular


@pytest.mark.gpu
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_stage_assembly_parity(stage):
    """Stage `stage` of DIRK-3,3 with a BDF2 history (the general formula):
    residual and Jacobian within 1e-12 of the oracle."""
    from mrhyde_amd import Mhx
    from mrhyde_amd.abi import butcher_tableau
    from oracle import Oracle

    p = TR_CONFIGS["2d_q2_nonlinear_perturbed"]
    g, o = Mhx(p), Oracle(p)
    A, b, _ = butcher_tableau("DIRK-3,3")
    rng = np.random.default_rng(40 + stage)
    hist = [rng.standard_normal(o.n_dof) for _ in range(2)]
    stages = [rng.standard_normal(o.n_dof) for _ in range(stage)]
    u = o.setDirichlet(rng.standard_normal(o.n_dof))
    print(f"[transient] DIRK-3,3 stage {stage}", flush=True)
    g.set_stage_state(0.05, 2, hist, A, b, stage, stages)
    o.set_stage_state(0.05, 2, hist, A, b, stage, stages)
    rg, jg = g.assembleJacRes(u)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"stage {stage} Jacobian")
    assert_entries_close(rg, ro, None, what=f"stage {stage} residual")
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tableau", ["DIRK-2,3", "CN"])
def test_stage_method_trajectory_matches_oracle(tableau):
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = TR_CONFIGS["2d_q2_nonlinear_perturbed"]
    g, o = Mhx(p), Oracle(p)
    traj_g, traj_o = [], []
    u0 = np.zeros(o.n_dof)
    g.transientSolver(u0, 0.05, 4, tableau=tableau, record=lambda t, u: traj_g.append(u.copy()))
    o.transientSolver(u0, 0.05, 4, tableau=tableau, record=lambda t, u: traj_o.append(u.copy()),
                      opts=newton_opts())
    for a, b in zip(traj_g, traj_o):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11)
    g.close()
