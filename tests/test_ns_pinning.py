"""Pins the Navier-Stokes restatement in the oracle (navierstokes.cpp:252-850,
computeTau :1053-1081) beyond the one reference gold that exercises it
(regression/navierstokes/channel, 2-D, PSPG only):

* rotated channel: the channel gold rotated so the flow runs along y or z of
  a 3-D box (the third axis extruded with the normal velocity held at 0): the
  L2 errors equal the 2-D gold values times sqrt(depth) for every
  orientation, which needs the z convection and the z pressure gradient;
* exactness: every affine flow u = c + a (b . x) with a orthogonal to b is
  divergence-free, has constant convection a (b . c), and with a linear
  pressure p = rho (f - a (b . c)) . x its strong momentum residual vanishes;
  the Q1-Q1 space holds it on any (perturbed) mesh, so the SUPG + PSPG
  discrete solution must equal it at every node;
* closed form: on one box element with constant velocity U and linear
  pressure, the SUPG and PSPG residuals reduce to tau R (U . int grad phi_i)
  and tau / rho R . int grad phi_i with int d_c phi_i = +-|K| / (2^(d-1) h_c);
* Kovasznay flow (an exact steady solution with no body force): measured
  velocity L2 convergence with SUPG + PSPG on, in 2-D, and the same flow
  rotated into the (z, x) plane of a one-layer 3-D box giving the 2-D errors
  times sqrt(depth).

All on the CPU oracle; Newton uses sparse LU (scipy) so the iterates do not
depend on a Krylov tolerance.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import GOLDEN, field_l2_errors, to_scipy
from mrhyde_amd.abi import MHX_PHYS_NAVIER_STOKES, Problem
from oracle import Oracle

NS_GOLD = json.load(open(os.path.join(GOLDEN, "ns_gold.json")))
C1, C2 = 4.0, 2.0  # computeTau constants (navierstokes.cpp:1057-1058)


def newton_lu(o, u, iters=15, tol=1e-13):
    """Newton with a direct solve; Dirichlet rows are e_d with residual 0, so
    whatever u holds on fixed DOFs stays there."""
    rp, ci = o.crs_pattern()
    r0 = None
    for _ in range(iters):
        r, j = o.assembleJacRes(u)
        nr = np.abs(r).max()
        r0 = nr if r0 is None else r0
        if nr <= tol * max(r0, 1.0):
            break
        u = u + spla.spsolve(to_scipy(rp, ci, j).tocsc(), r)
    return u


def node_field(o, nf):
    """(field of each DOF, coordinates of each DOF)."""
    x = o.dof_coords()
    return np.arange(o.n_dof) % nf, x


# --- rotated channel gold -----------------------------------------------------------


def channel_rotated(deck, flow, wall, ext, depth):
    """The channel deck with the flow along axis `flow`, walls normal to `wall`
    and `ext` extruded (2 element layers of `depth`/2, normal velocity 0 on its
    faces, natural conditions for the other components)."""
    n, hi = [0, 0, 0], [0.0, 0.0, 0.0]
    n[flow], n[wall], n[ext] = deck["n"][0], deck["n"][1], 2
    hi[flow], hi[wall], hi[ext] = deck["hi"][0], deck["hi"][1], depth
    force = [0.0, 0.0, 0.0]
    force[flow] = deck["force"][0]
    comp = [0, 0, 0]
    comp[ext] = 3 << (2 * ext)
    return Problem(dim=3, order=1, quadrature=deck["quadrature"], n=n, lo=(0.0, 0.0, 0.0), hi=hi,
                   physics=MHX_PHYS_NAVIER_STOKES, dbc_sides=3 << (2 * wall), rho=deck["rho"], mu=deck["mu"],
                   pspg=deck["pspg"], supg=deck["supg"], force=force, comp_sides=comp)


@pytest.mark.parametrize("flow,wall,ext", [(1, 2, 0), (2, 0, 1)], ids=["flow_y", "flow_z"])
def test_ns_rotated_channel_gold(flow, wall, ext):
    deck = NS_GOLD["navierstokes/channel"]
    depth = 0.2  # two layers of the 2-D element size 0.1, so h and tau are unchanged
    o = Oracle(channel_rotated(deck, flow, wall, ext, depth))
    u = newton_lu(o, o.setDirichlet(np.zeros(o.n_dof)))
    exact = [lambda x: 0.0 * x[:, 0]] * 4
    exact[flow] = lambda x: 0.5 * x[:, wall] * (1 - x[:, wall])
    e = field_l2_errors(o, u, exact)
    gold = deck["gold"]
    s = np.sqrt(depth)
    for f, name in [(flow, "ux"), (wall, "uy"), (3, "pr")]:
        assert abs(e[f] - s * gold[name]) <= 5e-6 * s * gold[name], (name, e[f], s * gold[name])
    assert e[ext] < 1e-10


# --- exactness for flows with zero strong residual ------------------------------------


def affine_flow(dim, seed, rho, force):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(dim)
    b = rng.standard_normal(dim)
    b -= a * (a @ b) / (a @ a)  # a . b = 0: trace of the velocity gradient a b^T
    c = rng.standard_normal(dim)
    f = np.asarray(force[:dim], dtype=float)
    gp = rho * (f - a * (b @ c))  # grad p = rho f - rho (u . grad) u
    vel = lambda x: c[None, :] + np.outer(x[:, :dim] @ b, a)
    pres = lambda x: x[:, :dim] @ gp + 0.3
    return vel, pres


@pytest.mark.parametrize("dim,n,seed", [(2, (5, 4), 1), (3, (3, 3, 2), 2), (3, (2, 3, 3), 3)])
@pytest.mark.parametrize("stab", ["supg+pspg", "pspg"])
def test_ns_exact_for_zero_strong_residual(dim, n, seed, stab):
    rho, mu = 1.3, 0.07
    force = (0.4, -0.3, 0.6)
    p = Problem(dim=dim, order=1, quadrature=2, n=n, physics=MHX_PHYS_NAVIER_STOKES, rho=rho, mu=mu,
                supg=stab == "supg+pspg", pspg=True, force=force, dbc_sides=(1 << (2 * dim)) - 1,
                pin_pressure=True, perturb=0.25, perturb_seed=seed)
    o = Oracle(p)
    nf = dim + 1
    fld, x = node_field(o, nf)
    vel, pres = affine_flow(dim, seed, rho, force)
    ex = np.where(fld == dim, pres(x), 0.0)
    v = vel(x)
    for a in range(dim):
        ex = np.where(fld == a, v[:, a], ex)
    mask, _ = o.dirichlet()
    u0 = np.where(mask == 1, ex, 0.0)
    # the exact field is a discrete solution: its residual vanishes
    assert np.abs(o.assembleRes(ex)).max() <= 1e-12 * max(1.0, np.abs(ex).max())
    u = newton_lu(o, u0)
    np.testing.assert_allclose(u, ex, rtol=0, atol=1e-10 * np.abs(ex).max())


# --- closed-form stabilisation terms on one box element --------------------------------


def box_residual(dim, h, U, g, force, rho, mu, supg, pspg):
    p = Problem(dim=dim, order=1, quadrature=2, n=(1,) * dim, lo=(0.0,) * 3, hi=tuple(h) + (1.0,) * (3 - dim),
                physics=MHX_PHYS_NAVIER_STOKES, rho=rho, mu=mu, supg=supg, pspg=pspg, force=force, dbc_sides=0)
    o = Oracle(p)
    nf = dim + 1
    fld, x = node_field(o, nf)
    u = np.where(fld == dim, x[:, :dim] @ np.asarray(g), 0.0)
    for a in range(dim):
        u = np.where(fld == a, U[a], u)
    return o.assembleRes(u), fld, x


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("speed", [1.7, 3e-7], ids=["u", "tiny_u"])
def test_ns_stabilisation_closed_form(dim, speed):
    rng = np.random.default_rng(40 + dim)
    h = [0.3, 0.2, 0.25][:dim]
    U = speed * rng.standard_normal(dim)
    g = rng.standard_normal(dim)
    force = list(rng.standard_normal(3))
    rho, mu = 0.9, 0.03
    vol = float(np.prod(h))
    he = vol ** (1.0 / dim)  # Workset::getElementSize
    n2 = float(U @ U)
    nvel = np.sqrt(n2) if n2 > 1e-12 else n2  # computeTau takes the root only above 1e-12
    tau = 1.0 / np.sqrt((C1 * mu / he ** 2) ** 2 + (C2 * nvel / he) ** 2)
    R = g - rho * np.asarray(force[:dim])  # rho (U . grad) U + grad p - rho f with constant U
    base, fld, x = box_residual(dim, h, U, g, force, rho, mu, False, False)
    on_s, _, _ = box_residual(dim, h, U, g, force, rho, mu, True, False)
    on_p, _, _ = box_residual(dim, h, U, g, force, rho, mu, False, True)
    nodes = x[::dim + 1, :dim]
    # int_K d_c phi_i = s_c(i) |K| / (2^(d-1) h_c), s = +1 on the high face of axis c
    s = np.where(nodes > 0.5 * np.asarray(h)[None, :], 1.0, -1.0)
    ig = s * vol / (2 ** (dim - 1) * np.asarray(h)[None, :])
    supg = np.zeros(len(fld))
    pspg = np.zeros(len(fld))
    for i in range(len(nodes)):
        for a in range(dim):
            supg[i * (dim + 1) + a] = tau * R[a] * (U @ ig[i])
        pspg[i * (dim + 1) + dim] = tau / rho * (R @ ig[i])
    # the residual returned is -F
    scale = max(np.abs(supg).max(), np.abs(pspg).max(), 1e-300)
    np.testing.assert_allclose(-(on_s - base), supg, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(-(on_p - base), pspg, rtol=0, atol=1e-12 * scale)


# --- Kovasznay flow -----------------------------------------------------------------------


def kovasznay(re):
    lam = re / 2 - np.sqrt(re * re / 4 + 4 * np.pi ** 2)
    u = lambda X, Y: 1 - np.exp(lam * X) * np.cos(2 * np.pi * Y)
    v = lambda X, Y: lam / (2 * np.pi) * np.exp(lam * X) * np.sin(2 * np.pi * Y)
    pr = lambda X, Y: -0.5 * np.exp(2 * lam * X)
    return u, v, pr


def kovasznay_solve(n, re, rotated=False):
    """Kovasznay flow on [0,1] x [-1/2,1/2] (n x n elements, rho = 1, mu = 1/re),
    SUPG + PSPG, exact velocity on the boundary, pressure pinned at the first
    node.  rotated: the 2-D axes (X, Y) become (z, x) of a 3-D box one element
    of h deep in y, uy = 0 on the y faces and natural conditions for ux, uz
    there.  Returns the L2 errors of (u_X, u_Y, p) (p up to nothing: the pin
    fixes its constant)."""
    uX, uY, pr = kovasznay(re)
    h = 1.0 / n
    if not rotated:
        p = Problem(dim=2, order=1, quadrature=2, n=(n, n), lo=(0.0, -0.5), hi=(1.0, 0.5),
                    physics=MHX_PHYS_NAVIER_STOKES, rho=1.0, mu=1.0 / re, supg=True, pspg=True, dbc_sides=15,
                    pin_pressure=True)
        XY = lambda x: (x[:, 0], x[:, 1])
        fX, fY, fp, nf = 0, 1, 2, 3
    else:
        p = Problem(dim=3, order=1, quadrature=2, n=(n, 1, n), lo=(-0.5, 0.0, 0.0), hi=(0.5, h, 1.0),
                    physics=MHX_PHYS_NAVIER_STOKES, rho=1.0, mu=1.0 / re, supg=True, pspg=True,
                    dbc_sides=1 | 2 | 16 | 32, comp_sides=(0, 4 | 8, 0), pin_pressure=True)
        XY = lambda x: (x[:, 2], x[:, 0])
        fX, fY, fp, nf = 2, 0, 3, 4
    o = Oracle(p)
    fld, x = node_field(o, nf)
    X, Y = XY(x)
    ex = np.zeros(o.n_dof)
    ex = np.where(fld == fX, uX(X, Y), ex)
    ex = np.where(fld == fY, uY(X, Y), ex)
    ex = np.where(fld == fp, pr(X, Y), ex)
    mask, _ = o.dirichlet()
    u = newton_lu(o, np.where(mask == 1, ex, 0.0))
    exact = [lambda x: 0.0 * x[:, 0]] * nf
    exact[fX] = lambda x: uX(*XY(x))
    exact[fY] = lambda x: uY(*XY(x))
    exact[fp] = lambda x: pr(*XY(x))
    e = field_l2_errors(o, u, exact)
    return e[fX], e[fY], e[fp], e


def test_ns_kovasznay_convergence():
    """Re = 40: the velocity L2 error of the SUPG + PSPG solution approaches
    second order (measured 1.74 and 1.85 on n = 16, 32, 64; 1.93 at 128) and
    the pressure first order (0.63, 0.82; 0.94 at 128), the orders rising
    with refinement."""
    errs = [kovasznay_solve(n, 40.0)[:3] for n in (16, 32, 64)]
    eu = [np.hypot(a, b) for a, b, _ in errs]
    ep = [c for _, _, c in errs]
    rates_u = [np.log2(eu[k] / eu[k + 1]) for k in range(2)]
    rates_p = [np.log2(ep[k] / ep[k + 1]) for k in range(2)]
    assert rates_u[1] >= 1.8 and rates_u[1] > rates_u[0] >= 1.65, (rates_u, eu)
    assert rates_p[1] >= 0.75 and rates_p[1] > rates_p[0], (rates_p, ep)
    # absolute level at n = 64 (a consistent method on this flow)
    assert eu[2] < 1.5e-3, eu


def test_ns_kovasznay_rotated_3d_equals_2d():
    """The 2-D Kovasznay solution of n = 12 and the same flow in the (z, x)
    plane of a 3-D box one element deep: L2 errors equal up to sqrt(depth),
    with the z convection, z pressure gradient and SUPG along z active."""
    n, re = 12, 40.0
    e2 = kovasznay_solve(n, re)
    e3 = kovasznay_solve(n, re, rotated=True)
    s = np.sqrt(1.0 / n)
    for a, b in zip(e2[:3], e3[:3]):
        assert abs(b - s * a) <= 1e-9 * s * a, (a, b)
    assert e3[3][1] < 1e-12  # uy


# --- traction (Neumann) sides, navierstokes.cpp:853-1000 ------------------------------
def _traction_problem(dim, flux, sides, perturb=0.0):
    n = (4, 3) if dim == 2 else (3, 3, 2)
    return Problem(dim=dim, order=1, quadrature=2, n=n, physics=MHX_PHYS_NAVIER_STOKES, mu=0.05, rho=1.3,
                   supg=True, pspg=True, dbc_sides=1, neumann_sides=sides, neumann_flux=flux, perturb=perturb,
                   perturb_seed=5)


@pytest.mark.parametrize("dim", [2, 3])
def test_ns_traction_load_integrates_the_traction(dim):
    """At u = 0 with no body force every volume term vanishes, so the
    residual is the traction load alone: summed over the loaded side's rows
    of component c it is the traction's integral over the side (partition of
    unity), constant or varying along the side, and the pressure rows and
    the other sides get nothing."""
    half_pi = np.pi / 2
    const = ([0.7, -0.4, 0.25][:dim] + [0.0] * (3 - dim), [0.0] * 3, [half_pi] * 3)  # sin(pi/2) = 1
    o = Oracle(_traction_problem(dim, [const], 2))  # the x = 1 side
    r = o.assembleRes(np.zeros(o.n_dof))
    nf = dim + 1
    coords = o.dof_coords()[::nf]
    on = np.isclose(coords[:, 0], 1.0)
    area = 1.0  # the unit box's x = 1 side
    for c in range(dim):
        np.testing.assert_allclose(r[c::nf][on].sum(), const[0][c] * area, rtol=1e-13)
        assert np.abs(r[c::nf][~on]).max() == 0.0
    assert np.abs(r[dim::nf]).max() == 0.0
    # a traction varying along y: the 2-point Gauss rule against the exact integral
    w = 1.3
    vary = ([1.0, 0.0, 0.0], [0.0, w, 0.0], [half_pi, 0.2, half_pi])  # q_x = sin(w y + 0.2)
    o2 = Oracle(_traction_problem(dim, [vary], 2))
    r2 = o2.assembleRes(np.zeros(o2.n_dof))
    exact = (np.cos(0.2) - np.cos(w + 0.2)) / w
    np.testing.assert_allclose(r2[0::nf][on].sum(), exact, rtol=1e-4)
