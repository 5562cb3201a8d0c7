"""Smoothed-aggregation AMG (MHX_PREC_AMG).

CPU: the oracle's hierarchy (oracle/amg.py, which follows the device
kernels' index arithmetic and checks every index) against scipy products of
the same operators, and its V-cycle as a PCG/GMRES preconditioner.
GPU: the device hierarchy against the oracle's (patterns bit-exact, values
within 1e-12 of the operator's largest entry), and Krylov solves with the AMG
V-cycle to 1e-10 true relative residual in FP64 and FP32-refined form.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mrhyde_amd.abi import MHX_PHYS_NAVIER_STOKES, Problem

PI = np.pi


def lognormal(n, s, seed):
    return np.exp(s * np.random.default_rng(seed).standard_normal(n))


CASES = {
    "2d_q1_linear": (Problem(dim=2, order=1, quadrature=2, n=(32, 24), modes=[(2 * PI * PI, (PI, PI))]), True),
    "3d_q1_nonlinear": (Problem(dim=3, order=1, quadrature=2, n=(10, 9, 8), diff_u2=1.0, k0=lognormal(720, 0.5, 1),
                                src_elem=np.ones(720), dbc_sides=1 | 2), False),
    "3d_q2_linear": (Problem(dim=3, order=2, quadrature=4, n=(6, 5, 5), k0=lognormal(150, 0.5, 2),
                             src_elem=np.ones(150), dbc_sides=1 | 2), True),
    "2d_q2_perturbed": (Problem(dim=2, order=2, quadrature=4, n=(20, 18), diff_u2=0.5, modes=[(3.0, (PI, 2 * PI))],
                                perturb=0.2, perturb_seed=3, dbc_sides=1 | 4), False),
    "2d_ns_cavity": (Problem(dim=2, order=1, quadrature=2, n=(16, 12), physics=MHX_PHYS_NAVIER_STOKES, mu=0.05,
                             supg=True, pspg=True, dbc_sides=1 | 2 | 4, lid_sides=8, lid_u=(1.0, 0.0),
                             pin_pressure=True), False),
    "3d_ns_cavity": (Problem(dim=3, order=1, quadrature=2, n=(8, 7, 6), physics=MHX_PHYS_NAVIER_STOKES, mu=0.01,
                             supg=True, pspg=True, dbc_sides=1 | 2 | 4 | 16 | 32, lid_sides=8, lid_u=(1.0, 0.0, 0.0),
                             pin_pressure=True), False),
}


def lattice(p):
    return [p.order * p.n[d] + 1 for d in range(p.dim)] + [1] * (3 - p.dim)


def system(p, linear, seed=3):
    from oracle import Oracle

    o = Oracle(p)
    # the 3-D cavity is linearised at the start of a lid-driven run (velocity
    # noise 1e-3): at u ~ N(0, 1) its Jacobian is so convective that neither
    # block Jacobi nor any V-cycle here makes GMRES(50) converge
    scale = 0.0 if linear else 1e-3 if (p.physics == MHX_PHYS_NAVIER_STOKES and p.dim == 3) else 1.0
    u = o.setDirichlet(np.random.default_rng(seed).standard_normal(o.n_dof) * scale)
    r, j = o.assembleJacRes(u)
    rp, ci = o.crs_pattern()
    fixed, _ = o.dirichlet()
    return o, r, j, sp.csr_matrix((j, ci, rp), shape=(o.n_dof, o.n_dof)), fixed.astype(bool)


_HIER = {}


def oracle_hierarchy(name):
    from oracle import amg as oa

    if name not in _HIER:
        p, linear = CASES[name]
        o, r, j, A, fixed = system(p, linear)
        _HIER[name] = (oa.hierarchy(A, lattice(p), p.n_field, p.order, fixed), r, j, A)
    return _HIER[name]


# ---- CPU: the oracle against the mathematics --------------------------------
@pytest.mark.parametrize("name", list(CASES))
def test_oracle_hierarchy_matches_scipy_products(name):
    levels, r, j, A = oracle_hierarchy(name)
    assert len(levels) >= 2 and levels[-1]["A"].shape[0] <= 128
    p = CASES[name][0]
    fixed = np.zeros(A.shape[0], bool)
    from oracle import Oracle

    fixed[:] = Oracle(p).dirichlet()[0].astype(bool)
    ns = np.where(fixed, 0.0, 1.0)
    for L in levels[:-1]:
        A_l, P, R, Ac, T = L["A"], L["P"], L["R"], L["Ac"], L["T"]
        # the tentative prolongator reproduces the near-null space
        np.testing.assert_allclose(T @ L["ns_c"], ns, rtol=0, atol=1e-14)
        # T has orthonormal columns per aggregate and field
        np.testing.assert_allclose((T.T @ T).toarray(), np.diag((L["nrm"] > 0).astype(float)), atol=1e-14)
        Pref = (sp.identity(A_l.shape[0]) - L["omega"] * sp.diags(L["dinv"]) @ A_l) @ T
        assert abs(P - Pref).max() <= 1e-14 * abs(Pref).max()
        assert abs(R - P.T).max() == 0.0
        Acref = R @ A_l @ P
        assert abs(Ac - Acref).max() <= 1e-14 * abs(Acref).max()
        # the power estimate lies below rho(D^-1 A), and close to it when
        # D^-1 A is similar to a symmetric matrix (on the stabilised
        # Navier-Stokes operators ten steps can stop well short of it)
        DA = (sp.diags(L["dinv"]) @ A_l).toarray()
        rho = np.abs(np.linalg.eigvals(DA)).max()
        if CASES[name][0].physics != MHX_PHYS_NAVIER_STOKES:
            assert 0.5 * rho <= L["lam"] <= rho * (1 + 1e-12)
        else:  # a Rayleigh quotient of a nonsymmetric matrix: bounded by its 2-norm
            assert 0.0 < L["lam"] <= np.linalg.norm(DA, 2) * (1 + 1e-12)
        ns = L["ns_c"]


@pytest.mark.parametrize("name", ["2d_q1_linear", "3d_q2_linear"])
def test_oracle_vcycle_pcg(name):
    """The V-cycle is a symmetric positive preconditioner on SPD problems: PCG
    reaches 1e-10 in a few tens of iterations, Jacobi-PCG needs many more."""
    from oracle import amg as oa

    levels, _, j, A = oracle_hierarchy(name)
    # a random right-hand side (the manufactured source of 2d_q1_linear is an
    # eigenvector of the uniform-grid operator, which Jacobi-PCG solves in one step)
    r = np.random.default_rng(4).standard_normal(A.shape[0])
    M = spla.LinearOperator(A.shape, matvec=lambda v: oa.vcycle(levels, v))
    it = [0, 0]
    x, info = spla.cg(A, r, M=M, rtol=1e-10, maxiter=300, callback=lambda xk: it.__setitem__(0, it[0] + 1))
    assert info == 0 and np.linalg.norm(r - A @ x) <= 1e-10 * np.linalg.norm(r)
    spla.cg(A, r, M=sp.diags(1.0 / A.diagonal()), rtol=1e-10, maxiter=5000,
            callback=lambda xk: it.__setitem__(1, it[1] + 1))
    assert it[0] <= 30 and it[0] * 3 < it[1], it


@pytest.mark.parametrize("start", ["zero", "warm"])
@pytest.mark.parametrize("lam_scale", [1.0, 0.8])
def test_oracle_chebyshev_is_degree2_residual_polynomial(start, lam_scale):
    """The Chebyshev(2) smoother (Ifpack2 defaults that the reference's MueLu
    list leaves unset, linearAlgebraInterface_solvers.hpp:208-253: eigenvalue
    ratio 7, boost 1.1) multiplies the error along each eigenvector of D^-1 A
    (eigenvalue t) by p(t) = T_2((theta - t) / delta) / T_2(theta / delta) on
    [lmin, lmax] = [1.1 lam / 7, 1.1 lam], T_2(s) = 2 s^2 - 1; from zero (the
    pre-smoother) and from a nonzero iterate (the post-smoother).  lam below
    the spectral radius (0.8) leaves eigenvalues above lmax, where |p|
    exceeds the minimax bound 1 / T_2(theta / delta)."""
    from oracle import Oracle
    from oracle import amg as oa

    from conftest import to_scipy

    p = Problem(dim=2, order=1, quadrature=2, n=(7, 6), k0=lognormal(42, 0.7, 5), src_elem=np.ones(42),
                dbc_sides=1 | 4, perturb=0.2, perturb_seed=6)
    o = Oracle(p)
    _, j = o.assembleJacRes(o.setDirichlet(np.zeros(o.n_dof)))
    rp, ci = o.crs_pattern()
    mask, _ = o.dirichlet()
    free = mask == 0
    A = to_scipy(rp, ci, j)[free][:, free].toarray()  # the SPD free block
    d = np.diag(A)
    s = 1.0 / np.sqrt(d)
    t, W = np.linalg.eigh(s[:, None] * A * s[None, :])  # D^-1/2 A D^-1/2 = W diag(t) W^T
    V = s[:, None] * W  # eigenvectors of D^-1 A
    lam = lam_scale * t.max()
    lmax = 1.1 * lam
    lmin = lmax / 7.0
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    T2 = lambda z: 2 * z * z - 1
    poly = T2((theta - t) / delta) / T2(theta / delta)
    As = sp.csr_matrix(A)
    rng = np.random.default_rng(7)
    xs = V @ rng.standard_normal(len(t))  # exact solution with every eigencomponent
    b = A @ xs
    x0 = None if start == "zero" else rng.standard_normal(len(t))
    e0 = xs if x0 is None else xs - x0
    x1 = oa.chebyshev(As, 1.0 / d, lam, b, x0)
    c0 = np.linalg.solve(V, e0)
    c1 = np.linalg.solve(V, xs - x1)
    np.testing.assert_allclose(c1, poly * c0, rtol=0, atol=1e-10 * np.abs(c0).max())
    bound = 1.0 / T2(theta / delta)  # the minimax value on [lmin, lmax]
    inside = (t >= lmin) & (t <= lmax)
    assert np.abs(poly[inside]).max() <= bound + 1e-12
    if lam_scale < 1.0:
        assert t.max() > lmax and np.abs(poly[t > lmax]).max() > bound


# ---- GPU: device hierarchy and solves ---------------------------------------
def _close(a, b, what, rel=1e-12):
    a, b = a.tocsr(), b.tocsr()
    a.sort_indices()
    b.sort_indices()
    np.testing.assert_array_equal(a.indptr, b.indptr, err_msg=what)
    np.testing.assert_array_equal(a.indices, b.indices, err_msg=what)
    scale = np.abs(b.data).max()
    err = np.abs(a.data - b.data).max()
    assert err <= rel * scale, f"{what}: max error {err:.3e} vs scale {scale:.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_device_hierarchy_parity(name):
    from mrhyde_amd import Mhx

    levels, r, j, A = oracle_hierarchy(name)
    g = Mhx(CASES[name][0])
    print(f"[amg] {name} n={A.shape[0]} levels={len(levels)}", flush=True)
    nl = g.buildAMGPreconditioner(j)
    assert nl == len(levels)
    for l, L in enumerate(levels):
        Ag, meta = g.amg_level(l, "A")
        assert (meta["rows"], meta["levels"]) == (L["A"].shape[0], len(levels))
        if l == 0:
            np.testing.assert_array_equal(Ag.indptr, L["A"].indptr)
            np.testing.assert_array_equal(Ag.indices, L["A"].indices)
        else:
            _close(Ag, L["A"], f"A_{l}")
        if l < len(levels) - 1:
            assert abs(meta["lambda"] - L["lam"]) <= 1e-12 * L["lam"], (meta["lambda"], L["lam"])
            _close(g.amg_level(l, "P")[0], L["P"], f"P_{l}")
            _close(g.amg_level(l, "R")[0], L["R"], f"R_{l}")
        else:
            Cinv, _ = g.amg_level(l, "Cinv")
            np.testing.assert_allclose(Cinv, L["Cinv"], rtol=0, atol=1e-10 * np.abs(L["Cinv"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("precond", ["amg", "amg-fp32"])
@pytest.mark.parametrize("name,solver", [("2d_q1_linear", "cg"), ("3d_q2_linear", "cg"), ("3d_q1_nonlinear", "gmres"),
                                         ("2d_q2_perturbed", "gmres"), ("2d_ns_cavity", "gmres")])
def test_amg_solve(name, solver, precond):
    from mrhyde_amd import Mhx, krylov_opts

    levels, _, j, A = oracle_hierarchy(name)
    r = np.random.default_rng(6).standard_normal(A.shape[0])
    g = Mhx(CASES[name][0])
    print(f"[amg-solve] {name} {solver} {precond}", flush=True)
    x, st = g.linearSolver(j, r, opts=krylov_opts(solver=solver, tol=1e-10, max_iters=2000, precond=precond))
    rel = np.linalg.norm(r - A @ x) / np.linalg.norm(r)
    assert st["rel_res"] <= 1e-10 and rel <= 1.5e-10, (st, rel)
    x_direct = spla.spsolve(A.tocsc(), r)
    np.testing.assert_allclose(x, x_direct, rtol=0, atol=1e-7 * np.abs(x_direct).max())
    # against Jacobi on the same system
    _, stj = g.linearSolver(j, r, opts=krylov_opts(solver=solver, tol=1e-10, max_iters=20000, precond="jacobi"))
    assert st["iters"] < stj["iters"], (st, stj)


@pytest.mark.gpu
def test_amg_hierarchy_is_reproducible():
    """Two set-ups of the same matrix give bit-identical operators (every
    device sum runs in a fixed order)."""
    from mrhyde_amd import Mhx

    levels, r, j, A = oracle_hierarchy("3d_q1_nonlinear")
    g = Mhx(CASES["3d_q1_nonlinear"][0])
    g.buildAMGPreconditioner(j)
    first = [g.amg_level(1, "A")[0].data.copy(), g.amg_level(0, "P")[0].data.copy()]
    g.buildAMGPreconditioner(j)
    second = [g.amg_level(1, "A")[0].data, g.amg_level(0, "P")[0].data]
    for a, b in zip(first, second):
        np.testing.assert_array_equal(a, b)


# ---- block smoothers ---------------------------------------------------------
SMOOTHERS = {1: "block_jacobi", 2: "block_gauss_seidel"}


@pytest.mark.parametrize("name", list(CASES))
def test_colouring_separates_coupled_nodes(name):
    """The multicolour sweep updates all nodes of a colour at once, so no two
    nodes of one colour may share a matrix entry: on every level the pattern's
    lattice reach must not exceed the level's reach (colour stride reach + 1),
    and the colours cover every node."""
    from oracle import amg as oa

    levels, _, _, _ = oracle_hierarchy(name)
    nf = CASES[name][0].n_field
    for L in levels[:-1]:
        N = L["N"]
        col, ncol = oa.node_colours(N, L["reach"])
        assert set(np.unique(col)) == set(range(ncol)) or col.size < ncol
        M = L["A"].tocoo()
        ni, nj = M.row // nf, M.col // nf
        off = ni != nj
        assert not np.any(col[ni[off]] == col[nj[off]]), "coupled nodes share a colour"


@pytest.mark.parametrize("name", ["2d_q1_linear", "3d_q2_linear", "2d_ns_cavity"])
def test_oracle_gauss_seidel_sweep_is_block_triangular_solve(name):
    """A forward multicolour sweep from x = 0 is x = (D + L)^-1 b in the colour
    order: D the nodal diagonal blocks, L the blocks of earlier colours.  The
    right side is formed with scipy's sparse LU on the permuted matrix, which
    shares no code with the sweep."""
    from oracle import amg as oa

    levels, _, _, _ = oracle_hierarchy(name)
    nf = CASES[name][0].n_field
    L = levels[0]
    A = L["A"].tocsr()
    col, _ = oa.node_colours(L["N"], L["reach"])
    rows_col = np.repeat(col, nf)
    b = np.random.default_rng(8).standard_normal(A.shape[0])
    x = oa.gs_sweep(A, oa.block_inverses(A, nf), nf, L["N"], L["reach"], b, np.zeros_like(b), True)
    Ac = A.tocoo()
    node_r, node_c = Ac.row // nf, Ac.col // nf
    keep = (rows_col[Ac.row] > rows_col[Ac.col]) | (node_r == node_c)
    M = sp.csr_matrix((Ac.data[keep], (Ac.row[keep], Ac.col[keep])), shape=A.shape)
    ref = spla.spsolve(M.tocsc(), b)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("smoother", list(SMOOTHERS.values()))
@pytest.mark.parametrize("name", ["2d_q1_linear", "3d_q2_linear"])
def test_oracle_block_smoother_vcycle_pcg(name, smoother):
    """Both block smoothers give a symmetric V-cycle (damped Jacobi sweeps on
    both sides; forward then backward Gauss-Seidel): PCG converges."""
    from oracle import amg as oa

    levels, _, _, A = oracle_hierarchy(name)
    r = np.random.default_rng(4).standard_normal(A.shape[0])
    M = spla.LinearOperator(A.shape, matvec=lambda v: oa.vcycle(levels, v, smoother, 1))
    x, info = spla.cg(A, r, M=M, rtol=1e-10, maxiter=100)
    assert info == 0 and np.linalg.norm(r - A @ x) <= 1e-10 * np.linalg.norm(r)


@pytest.mark.gpu
@pytest.mark.parametrize("sm", list(SMOOTHERS))
@pytest.mark.parametrize("name", ["2d_q1_linear", "3d_q2_linear", "2d_ns_cavity", "3d_ns_cavity"])
def test_device_block_smoother_vcycle_parity(name, sm):
    """One GMRES iteration with the V-cycle as right preconditioner returns
    x = c z, z = M^-1 r and c = (Az . r) / (Az . Az): the device cycle against
    the oracle's."""
    from mrhyde_amd import Mhx, krylov_opts
    from oracle import amg as oa

    levels, _, j, A = oracle_hierarchy(name)
    nf = CASES[name][0].n_field
    r = np.random.default_rng(9).standard_normal(A.shape[0])
    g = Mhx(CASES[name][0])
    g.set_option("amg.smoother", sm)
    assert g.get_option("amg.smoother") == sm
    print(f"[amg-smoother] {name} {SMOOTHERS[sm]}", flush=True)
    x, st = g.linearSolver(j, r, opts=krylov_opts(solver="gmres", tol=1e-30, max_iters=1, precond="amg"))
    assert st["iters"] == 1
    z = oa.vcycle(levels, r, SMOOTHERS[sm], nf)
    w = A @ z
    ref = (w @ r) / (w @ w) * z
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("sm", list(SMOOTHERS))
@pytest.mark.parametrize("name,solver", [("3d_q2_linear", "cg"), ("2d_ns_cavity", "gmres"),
                                         ("3d_q1_nonlinear", "gmres"), ("3d_ns_cavity", "gmres")])
def test_block_smoother_amg_solve(name, solver, sm):
    from mrhyde_amd import Mhx, krylov_opts

    levels, _, j, A = oracle_hierarchy(name)
    r = np.random.default_rng(6).standard_normal(A.shape[0])
    g = Mhx(CASES[name][0])
    g.set_option("amg.smoother", sm)
    print(f"[amg-smoother-solve] {name} {solver} {SMOOTHERS[sm]}", flush=True)
    for prec in ("amg", "amg-fp32"):
        x, st = g.linearSolver(j, r, opts=krylov_opts(solver=solver, tol=1e-10, max_iters=500, precond=prec))
        rel = np.linalg.norm(r - A @ x) / np.linalg.norm(r)
        assert st["rel_res"] <= 1e-10 and rel <= 1.5e-10, (prec, st, rel)


@pytest.mark.gpu
def test_ns_newton_with_block_jacobi_amg():
    """Newton on a 3-D lid-driven cavity (Re 100) with GMRES(50) and the AMG
    cycle smoothed by nodal block Jacobi: every linear solve reaches 1e-10 and
    the solution equals a Newton iteration on the oracle's residual and
    Jacobian with sparse LU solves."""
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = Problem(dim=3, order=1, quadrature=2, n=(10, 10, 10), physics=MHX_PHYS_NAVIER_STOKES, mu=0.01, supg=True,
                pspg=True, dbc_sides=1 | 2 | 4 | 16 | 32, lid_sides=8, lid_u=(1.0, 0.0, 0.0), pin_pressure=True)
    o = Oracle(p)
    u0 = o.setDirichlet(1e-3 * np.random.default_rng(3).standard_normal(o.n_dof))
    rp, ci = o.crs_pattern()
    u_ref = u0.copy()
    for _ in range(6):
        r, j = o.assembleJacRes(u_ref)
        u_ref = u_ref + spla.spsolve(sp.csr_matrix((j, ci, rp), shape=(o.n_dof, o.n_dof)).tocsc(), r)
    g = Mhx(p)
    g.set_option("amg.smoother", 1)
    # also with the hierarchy kept across the Newton steps (fine-level block
    # inverses refreshed) and with the FP32 cycle in FP64 refinement
    for reuse, prec in ((0, "amg"), (1, "amg"), (1, "amg-fp32")):
        g.set_option("amg.reuse", reuse)
        print(f"[amg-newton] 10^3 cavity reuse={reuse} {prec}", flush=True)
        u, st = g.nonlinearSolver(u0.copy(), max_iters=10, nl_tol=1e-12, solver="gmres", restart=50, tol=1e-10,
                                  linear_max_iters=500, precond=prec)
        assert st["rel_res"] <= 1e-10 and st["res_norm"] <= 1e-12 * st["res_norm0"], (reuse, prec, st)
        assert st["iters"] <= 60 * st["nl_iters"], (reuse, prec, st)
        np.testing.assert_allclose(u, u_ref, rtol=0, atol=1e-9 * np.abs(u_ref).max())


@pytest.mark.gpu
def test_amg_options_are_validated():
    """The AMG options keep their defaults until set, accept their documented
    ranges and refuse anything else with MHX_ERR_ARG and a message."""
    from mrhyde_amd import Mhx
    from mrhyde_amd.engine import MhxError

    g = Mhx(CASES["2d_q1_linear"][0])
    assert (g.get_option("amg.smoother"), g.get_option("amg.sweeps"), g.get_option("amg.omega_pct")) == (-1, 3, 40)
    for name, good, bad in (("amg.smoother", (-1, 0, 1, 2), (-2, 3)), ("amg.sweeps", (1, 16), (0, 17)),
                            ("amg.omega_pct", (1, 200), (0, 201))):
        for v in good:
            g.set_option(name, v)
            assert g.get_option(name) == v
        for v in bad:
            with pytest.raises(MhxError, match=name.split(".")[1]):
                g.set_option(name, v)
            assert g.get_option(name) == good[-1]


@pytest.mark.gpu
@pytest.mark.parametrize("precond", ["amg", "amg-fp32"])
def test_amg_reuse_newton(precond):
    """amg.reuse keeps the coarse levels of the first Newton step's hierarchy
    and refreshes the fine-level smoother data from each new Jacobian: every
    linear solve still reaches 1e-10 and Newton lands on the same solution."""
    from mrhyde_amd import Mhx

    p = Problem(dim=3, order=2, quadrature=4, n=(8, 7, 6), diff_u2=1.0, k0=lognormal(336, 0.5, 21),
                src_elem=np.random.default_rng(22).uniform(0, 10, 336), dbc_sides=1 | 2)
    g = Mhx(p)
    sols = []
    for reuse in (0, 1):
        g.set_option("amg.reuse", reuse)
        u = g.setDirichlet(np.zeros(g.n_dof))
        u, st = g.nonlinearSolver(u, max_iters=8, nl_tol=1e-12, solver="gmres", restart=50, tol=1e-10,
                                  linear_max_iters=500, precond=precond)
        assert st["rel_res"] <= 1e-10 and st["res_norm"] <= 1e-12 * st["res_norm0"], (reuse, st)
        sols.append(u)
    np.testing.assert_allclose(sols[1], sols[0], rtol=0, atol=1e-9 * np.abs(sols[0]).max())
