"""Coefficient functions on the GPU: the compiled expressions evaluated in the
assembly kernel's point stage.

* Residual and Jacobian within 1e-12 of the oracle's FunctionManager
  restatement, steady and transient, for nonlinear diffusion, sources that
  depend on the solution, advection, element data and time.
* The expressions of the closed forms give the closed forms' bits:
  "k0*(1+T^2)" and "2*pi*pi*sin(pi*x)*sin(pi*y)" against diff_u2 = 1 and a
  sine mode, in the same kernel family.
* The reference's 2D_verification_transient deck through its own source
  string (with "t") reproduces the gold of every step.
* A context refuses what it cannot compile, with the message.
"""
import numpy as np
import pytest

from conftest import assert_entries_close
from mrhyde_amd.abi import Problem

pytestmark = pytest.mark.gpu
PI = np.pi


def lognormal(n, s, seed):
    return np.exp(s * np.random.default_rng(seed).standard_normal(n))


CONFIGS = {
    "2d_q1_c1_source": Problem(dim=2, order=1, quadrature=2, n=(16, 16),
                               functions={"thermal source": "2*pi*pi*sin(pi*x)*sin(pi*y)"}),
    "3d_q2_k0_nonlinear": Problem(dim=3, order=2, quadrature=4, n=(5, 4, 3), k0=lognormal(60, 0.5, 6),
                                  src_elem=np.random.default_rng(7).uniform(0, 1, 60), dbc_sides=1 | 2,
                                  functions={"thermal diffusion": "k0*(1+T^2)", "thermal source": "src"}),
    "2d_q3_perturbed_rich": Problem(
        dim=2, order=3, quadrature=6, n=(4, 3), perturb=0.2, perturb_seed=4, dbc_sides=1 | 4,
        functions={"thermal diffusion": "kk*(1+x*y)+max(T,0.2)", "kk": "exp(-T)",
                   "thermal source": "sin(pi*x)*T^2-abs(y-0.5)+sqrt(grad(T)[x]^2+grad(T)[y]^2)",
                   "advection x": "1+T", "advection y": "x-0.5*t", }, time=0.75),
    "3d_q1_perturbed_adv": Problem(
        dim=3, order=1, quadrature=2, n=(4, 3, 3), perturb=0.2, perturb_seed=8, dbc_sides=1 | 2 | 16,
        functions={"thermal diffusion": "1+0.5*T^2.5", "thermal source": "10*exp(-(x-0.5)^2-(z-0.5)^2)",
                   "advection z": "mean(T,1)"}),
}


def _state(o, seed):
    return o.setDirichlet(np.random.default_rng(seed).uniform(0.2, 1.0, o.n_dof))


@pytest.mark.parametrize("name", list(CONFIGS))
def test_function_assembly_parity(name):
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = CONFIGS[name]
    print(f"[functions] {name}", flush=True)
    g, o = Mhx(p), Oracle(p)
    assert g.get_option("assembly.colouring") == 0  # compiled functions run element colouring
    u = _state(o, 3)
    rg, jg = g.assembleJacRes(u)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"{name} Jacobian")
    assert_entries_close(rg, ro, None, what=f"{name} residual")
    # the residual-only build (its own instantiation: contraction may differ by an ulp)
    assert_entries_close(g.assembleRes(u), ro, None, what=f"{name} residual-only")
    g.close()


@pytest.mark.parametrize("order", [1, 2])
def test_function_transient_parity(order):
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = Problem(dim=2, order=2, quadrature=4, n=(5, 4), perturb=0.15, perturb_seed=2, dbc_sides=1 | 2, time=0.3,
                functions={"specific heat": "2.0+T*T", "density": "1.5+0*x",
                           "thermal source": "(8*(pi*pi)*sin(2*pi*t)+2*pi*cos(2*pi*t))*sin(2*pi*x)*sin(2*pi*y)",
                           "thermal diffusion": "1+T_t*0.1"})
    g, o = Mhx(p), Oracle(p)
    rng = np.random.default_rng(11 + order)
    hist = [rng.standard_normal(o.n_dof) for _ in range(order)]
    u = o.setDirichlet(rng.standard_normal(o.n_dof))
    print(f"[functions] transient order={order}", flush=True)
    g.set_time_state(0.05, order, hist)
    o.set_time_state(0.05, order, hist)
    rg, jg = g.assembleJacRes(u)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what="transient Jacobian")
    assert_entries_close(rg, ro, None, what="transient residual")
    g.close()


def test_functions_equal_closed_forms_bitwise():
    """Expressions that spell the closed forms give the same bits."""
    from mrhyde_amd import Mhx

    rng = np.random.default_rng(5)
    k0 = lognormal(60, 0.5, 9)
    base = dict(dim=3, order=2, quadrature=4, n=(5, 4, 3), k0=k0, dbc_sides=1 | 2)
    closed = Mhx(Problem(**base, diff_u2=1.0, modes=[(2 * PI * PI, (PI, PI, 0.0))]))
    closed.set_option("assembly.colouring", 0)  # the same kernel family as the compiled functions
    expr = Mhx(Problem(**base, functions={"thermal diffusion": "k0*(1+T^2)",
                                          "thermal source": "2*pi*pi*sin(pi*x)*sin(pi*y)"}))
    u = closed.setDirichlet(rng.standard_normal(closed.n_dof))
    rc, jc = closed.assembleJacRes(u)
    re, je = expr.assembleJacRes(u)
    np.testing.assert_array_equal(re, rc)
    np.testing.assert_array_equal(je, jc)
    closed.close()
    expr.close()
    # the C1 source on a 2-D Q1 mesh
    b2 = dict(dim=2, order=1, quadrature=2, n=(16, 16))
    c2 = Mhx(Problem(**b2, modes=[(2 * PI * PI, (PI, PI))]))
    e2 = Mhx(Problem(**b2, functions={"thermal source": "2*pi*pi*sin(pi*x)*sin(pi*y)"}))
    u2 = c2.setDirichlet(rng.standard_normal(c2.n_dof))
    r1, j1 = c2.assembleJacRes(u2)
    r2, j2 = e2.assembleJacRes(u2)
    np.testing.assert_array_equal(r2, r1)
    np.testing.assert_array_equal(j2, j1)
    c2.close()
    e2.close()


def test_transient_deck_string_reproduces_gold():
    from mrhyde_amd import Mhx
    from test_transient import GOLD, check_gold, gold_true

    fns = GOLD["functions"]
    assert "t" in fns["thermal source"]

    def at_time(t):
        return Problem(dim=GOLD["dim"], order=GOLD["order"], quadrature=GOLD["quadrature"], n=tuple(GOLD["n"]),
                       workset_size=GOLD["workset"], functions=fns, time=t)

    g = Mhx(at_time(0.0))
    errs = []
    dt = GOLD["final_time"] / GOLD["steps"]
    g.transientSolver(np.zeros(g.n_dof), dt, GOLD["steps"], order=GOLD["bdf_order"], at_time=at_time,
                      record=lambda t, u: errs.append((t, g.computeError(u, gold_true(t))[0])))
    check_gold(errs)
    g.close()


@pytest.mark.parametrize("fns,what", [
    ({"thermal source": "q+1"}, "could not decompose"),
    ({"thermal diffusion": "emax(T)"}, "element reductions"),
    ({"thermal diffusion": "a", "a": "b", "b": "a"}, "cyclic"),
])
def test_context_refuses_bad_functions(fns, what):
    from mrhyde_amd import Mhx, MhxError

    with pytest.raises(MhxError, match=what):
        Mhx(Problem(dim=2, order=1, quadrature=2, n=(3, 3), functions=fns))
