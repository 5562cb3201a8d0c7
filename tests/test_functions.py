"""Coefficient functions (the reference's FunctionManager / Interpreter) on
the CPU.

* The oracle's restatement is pinned against values worked out by hand from
  the reference's parsing and evaluation rules (interpreter.cpp:63-452,
  functionManager_evaluate.hpp:237-1353): precedence and left folds, the
  quirks of the split ('=' dropped inside sums, "2e-3" split at its '-'),
  the power rule for scalar exponents, sqrt of non-positive values, the
  element reductions (emax keeps the last point above the first, emean
  counts the first point twice), derivatives by the chain rule.
* The reference decks' own "thermal source" strings reproduce their gold L2
  errors (regression/thermal/*/mrhyde.gold) through the oracle.
* The product's compiler and evaluator (mhx_expr_eval, host) agree with the
  oracle on a seeded corpus of expressions, and refuse what the reference
  cannot evaluate.
* Expressions equal to the closed forms give the closed forms' assembly.
"""
import json
import os

import numpy as np
import pytest

from mrhyde_amd.abi import MHX_SIDE_ALL, Problem

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PI = np.pi


def pts_of(T=0.5, Tt=0.0, g=(0.0, 0.0, 0.0), x=(0.0, 0.0, 0.0), t=0.0, k0=1.0, src=0.0):
    return np.array([[T, Tt, *g, *x, t, k0, src]], dtype=np.float64)


def ev(expr, fns=None, **kw):
    import oracle

    fns = dict(fns or {})
    fns["f"] = expr
    return oracle.expr_eval(Problem(functions=fns), "f", pts_of(**kw))[0]


# --- oracle pinned by hand ----------------------------------------------------------
@pytest.mark.parametrize("expr,kw,val", [
    ("2-3+4", {}, 3.0),                    # left fold of + and -
    ("8/2/2", {}, 2.0),                    # left fold of /
    ("2^3^2", {}, 64.0),                   # left fold of ^: (2^3)^2, not 2^9
    ("-2^2", {}, -4.0),                    # "0.0-2^2"
    ("2*3+4*5", {}, 26.0),                 # * binds tighter than +
    ("1+2<3", {}, 2.0),                    # < binds tighter than +: 1 + (2<3)
    ("2*1e-3", {}, -1.0),                  # split at the '-' of 1e-3: (2*1e) - 3
    ("1e-3", {}, 1e-3),                    # a number alone
    ("(x<=0.4)", {"x": (0.4, 0, 0)}, 1.0),   # <= at the product level
    ("1+(x<=0.4)", {"x": (0.4, 0, 0)}, 1.0),  # '=' dropped inside a sum: x < 0.4
    ("x>=0.4", {"x": (0.4, 0, 0)}, 1.0),
    ("x>0.4", {"x": (0.4, 0, 0)}, 0.0),
    ("sqrt(-4)", {}, 0.0),                 # sqrt of a non-positive value is 0
    ("sqrt(x)", {"x": (0.25, 0, 0)}, 0.5),
    ("abs(x-1)", {"x": (0.25, 0, 0)}, 0.75),
    ("max(x,0.3)", {"x": (0.25, 0, 0)}, 0.3),
    ("min(x,0.3)", {"x": (0.25, 0, 0)}, 0.25),
    ("mean(x,0.75)", {"x": (0.25, 0, 0)}, 0.5),
    ("2*pi", {}, 2 * PI),
    ("sin(pi*(x+1))", {"x": (0.5, 0, 0)}, np.sin(PI * 1.5)),  # parentheses inside a one-argument operator
    ("k0*src+t", {"k0": 3.0, "src": 2.0, "t": 0.5}, 6.5),
    ("u+v+w+1", {}, 1.0),                  # mesh displacement fields: zero
    ("sinh(1)+cosh(1)", {}, np.sinh(1.0) + np.cosh(1.0)),
    ("exp(log(3))", {}, np.exp(np.log(3.0))),
    ("tan(0.25)", {}, np.tan(0.25)),
    ("emax(2.0)+emin(t)+emean(3)", {"t": 1.5}, 6.5),   # reductions of scalars: themselves
    ("a*2", {"fns": {"a": "x+1"}, "x": (0.5, 0, 0)}, 3.0),  # another function by name
])
def test_oracle_values_by_hand(expr, kw, val):
    fns = kw.pop("fns", None)
    assert ev(expr, fns, **kw)[0] == pytest.approx(val, rel=1e-15, abs=1e-300)


@pytest.mark.parametrize("expr,T,dT", [
    ("T^2", 0.7, 2 * 0.7),                          # scalar exponent 2: T*T
    ("T^3", 0.7, 3 * 0.7 ** 2),
    ("T^2.5", 0.7, 2.5 * 0.7 ** 1.5),               # pow rule
    ("T^2.5", 0.0, 0.0),                            # no derivative at a zero base
    ("2^T", 0.7, np.log(2) * 2 ** 0.7),
    ("sin(T)*T", 0.7, np.cos(0.7) * 0.7 + np.sin(0.7)),
    ("exp(-T)", 0.7, -np.exp(-0.7)),
    ("1/(1+T)", 0.7, -1 / 1.7 ** 2),
    ("abs(T-1)", 0.7, -1.0),
    ("sqrt(T)", 0.7, 0.5 / np.sqrt(0.7)),
    ("max(T,0.5)", 0.7, 1.0),
    ("max(T,0.9)", 0.7, 0.0),
    ("mean(T,T*T)", 0.7, 0.5 + 0.7),
    ("(T<0.9)*T", 0.7, 1.0),                        # comparisons have no derivative
    ("log(T)", 0.7, 1 / 0.7),
    ("tan(T)", 0.7, 1 + np.tan(0.7) ** 2),
    ("sinh(T)-cosh(T)", 0.7, np.cosh(0.7) - np.sinh(0.7)),
])
def test_oracle_derivatives_by_hand(expr, T, dT):
    r = ev(expr, T=T)
    assert r[1] == pytest.approx(dT, rel=1e-13, abs=1e-300)
    assert np.all(r[2:] == 0.0)


def test_oracle_gradient_and_time_slots():
    # d/dgrad(T)[c] and d/dT_t land in their own slots
    r = ev("grad(T)[x]*grad(T)[x]+3*grad(T)[y]-grad(T)[z]+2*T_t", g=(0.5, 1.0, 2.0), Tt=0.25)
    assert r[0] == pytest.approx(0.25 + 3.0 - 2.0 + 0.5)
    np.testing.assert_array_equal(r[1:], [0.0, 1.0, 3.0, -1.0, 2.0])


def test_oracle_element_reductions():
    import oracle

    # one element of three points with x = 1, 3, 2
    pts = np.zeros((3, 11))
    pts[:, 5] = [1.0, 3.0, 2.0]
    p = Problem(functions={"a": "emax(x)", "b": "emin(x)", "c": "emean(x)"})
    # emax: the last point whose value exceeds the first point's
    np.testing.assert_array_equal(oracle.expr_eval(p, "a", pts, one_element=True)[:, 0], [2.0, 2.0, 2.0])
    # emin: no point below the first -> the first
    np.testing.assert_array_equal(oracle.expr_eval(p, "b", pts, one_element=True)[:, 0], [1.0, 1.0, 1.0])
    # emean: t0/n + sum_q t_q/n (the first point counted twice)
    np.testing.assert_allclose(oracle.expr_eval(p, "c", pts, one_element=True)[:, 0], [7.0 / 3.0] * 3, rtol=1e-15)


@pytest.mark.parametrize("expr", [
    "x*-1",              # "x*" minus 1: an empty operand
    "q+1",               # unknown name
    "sin(x",             # unclosed parenthesis
    "x+thermal diffusion",  # spaces dropped in the sum: "thermaldiffusion" is no function
])
def test_oracle_refuses(expr):
    with pytest.raises(RuntimeError):
        ev(expr, {"thermal diffusion": "1.0"})


def test_oracle_cycle_refused():
    import oracle

    with pytest.raises(RuntimeError):
        oracle.expr_eval(Problem(functions={"a": "b+1", "b": "a*2"}), "a", pts_of())


# --- the reference decks' expressions reproduce their gold --------------------------
GOLD = json.load(open(os.path.join(GOLDEN, "thermal_gold.json")))
CASES = [k for k in GOLD if not k.startswith("_")]


def deck_volume_functions(g):
    # functions at volume points (side functions such as "thermal Neumann
    # source" live at "side ip" in the reference)
    return {k: v for k, v in g.get("functions", {}).items() if "Neumann" not in k}


@pytest.mark.parametrize("case", CASES)
def test_deck_expressions_reproduce_gold(case):
    from conftest import thermal_gold_problem
    from oracle import Oracle
    from test_oracle import nopts

    p, true, g = thermal_gold_problem(GOLD[case])
    fns = deck_volume_functions(g)
    assert "thermal source" in fns
    p.modes = ()
    p.functions = fns
    o = Oracle(p)
    u = o.setDirichlet(np.zeros(o.n_dof))
    u, _ = o.nonlinearSolver(u, nopts(max_iters=g["nl_iters"], nl_tol=g["nl_tol"]))
    l2, _ = o.computeError(u, true)
    assert float(f"{l2:.6g}") == g["gold"]["T"], (l2, g["gold"]["T"])


# --- product compiler + evaluator (host) against the oracle --------------------------
def random_expr(rng, depth=0):
    leaves = ["T", "grad(T)[x]", "grad(T)[y]", "grad(T)[z]", "T_t", "x", "y", "z", "t", "pi", "k0", "src",
              "0.5", "2", "3", "1.25"]
    if depth > 3 or rng.random() < 0.3:
        return str(rng.choice(leaves))
    kind = rng.integers(0, 6)
    a, b = random_expr(rng, depth + 1), random_expr(rng, depth + 1)
    if kind == 0:
        return f"{a}{rng.choice(['+', '-', '*', '/'])}{b}"
    if kind == 1:
        return f"({a}){rng.choice(['+', '-', '*', '/'])}({b})"
    if kind == 2:
        fn = rng.choice(["sin", "cos", "exp", "sqrt", "abs", "sinh", "cosh", "tan"])
        return f"{fn}({a})"
    if kind == 3:
        return f"({a})^{rng.choice(['2', '3', '4', '2.5', '0.5', '2.0000000000001'])}"
    if kind == 4:
        # two-argument operators only with parenthesis-free arguments (the
        # reference's operator form)
        a, b = str(rng.choice(leaves)), str(rng.choice(leaves))
        return f"{rng.choice(['max', 'min', 'mean'])}({a},{b})"
    return f"({a}){rng.choice(['<', '>', '<=', '>='])}({b})"


def test_product_matches_oracle_on_corpus():
    import oracle
    from mrhyde_amd import function_eval

    rng = np.random.default_rng(20261203)
    pts = rng.uniform(0.1, 0.9, (64, 11))
    checked = 0
    for i in range(300):
        expr = random_expr(rng)
        p = Problem(functions={"f": expr})
        try:
            want = oracle.expr_eval(p, "f", pts)
        except RuntimeError:
            with pytest.raises(Exception):
                function_eval(p, "f", pts)
            continue
        got = function_eval(p, "f", pts)
        scale = np.maximum(np.abs(want), 1.0)
        with np.errstate(invalid="ignore"):
            ok = (got == want) | (np.isnan(want) & np.isnan(got)) | (np.abs(got - want) <= 1e-13 * scale)
        assert np.all(ok), (expr, got[~ok][:4], want[~ok][:4])
        checked += 1
    assert checked > 250


@pytest.mark.parametrize("expr,why", [
    ("emax(T)", "element reduction of a field"),
    ("h*2", "h"),
    ("foo(x)+1", "unknown operator"),
    ("max(x,(y))", "parentheses inside a two-argument operator"),
    ("sin(x,y)", "argument count"),
    ("x*", "empty operand"),
])
def test_product_refuses(expr, why):
    from mrhyde_amd import MhxError, function_eval

    with pytest.raises(MhxError):
        function_eval(Problem(functions={"f": expr}), "f", pts_of())


# --- expressions equal to the closed forms -------------------------------------------
def test_oracle_expressions_equal_closed_forms():
    from oracle import Oracle

    rng = np.random.default_rng(5)
    n = (4, 3)
    k0 = np.exp(0.5 * rng.standard_normal(12))
    base = dict(dim=2, order=2, quadrature=4, n=n, k0=k0, dbc_sides=MHX_SIDE_ALL & 3)
    closed = Problem(**base, diff_u2=1.0, modes=[(2 * PI * PI, (PI, PI))])
    expr = Problem(**base, functions={"thermal diffusion": "k0*(1+T^2)",
                                      "thermal source": "2*pi*pi*sin(pi*x)*sin(pi*y)"})
    oc, oe = Oracle(closed), Oracle(expr)
    u = oc.setDirichlet(rng.standard_normal(oc.n_dof))
    rc, jc = oc.assembleJacRes(u)
    re, je = oe.assembleJacRes(u)
    np.testing.assert_array_equal(re, rc)
    np.testing.assert_array_equal(je, jc)
