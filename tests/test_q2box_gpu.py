"""The pencil-streamed Q2 box assembly (assembly.colouring = 3) against the
oracle: every Jacobian entry and residual within 1e-12, for x-segments of
several lengths (segment ends inside the mesh, odd tails with one element
in the last step), accumulate mode, a sine-mode source, Dirichlet rows on
every side, and the same bits on a second assembly and against the
residual-only pass.  Meshes with ragged extents (1, 2, 3, 5, 7 elements per
axis) cover the pencil and segment boundaries on all three axes."""
import numpy as np
import pytest

from conftest import assert_entries_close

pytestmark = pytest.mark.gpu

PI = np.pi


def lognormal(n, s, seed):
    return np.exp(s * np.random.default_rng(seed).standard_normal(n))


def _configs():
    from mrhyde_amd.abi import Problem

    c = {}
    c["c2shape_5x4x3"] = Problem(dim=3, order=2, quadrature=4, n=(5, 4, 3), diff_u2=1.0, k0=lognormal(60, 0.5, 6),
                                 src_elem=np.random.default_rng(7).uniform(0, 1, 60), dbc_sides=1 | 2)
    c["modes_7x3x2_all_dbc"] = Problem(dim=3, order=2, quadrature=4, n=(7, 3, 2), diff_u2=0.7,
                                       k0=lognormal(42, 0.5, 21),
                                       modes=[(2.0, (PI, PI, PI)), (0.5, (2 * PI, 0, 3 * PI))],
                                       lo=(-0.5, 0.0, 1.0), hi=(1.5, 0.5, 1.75))
    c["line_8x1x1"] = Problem(dim=3, order=2, quadrature=4, n=(8, 1, 1), diff_u2=1.0, k0=lognormal(8, 1.0, 22),
                              src_elem=np.linspace(-1, 1, 8), dbc_sides=1)
    c["single_1x1x1"] = Problem(dim=3, order=2, quadrature=4, n=(1, 1, 1), diff_u2=2.0, src_elem=np.ones(1),
                                dbc_sides=0)
    c["plate_2x5x7"] = Problem(dim=3, order=2, quadrature=4, n=(2, 5, 7), diff_u2=0.3, k0=lognormal(70, 0.5, 23),
                               src_elem=np.random.default_rng(24).uniform(0, 1, 70), dbc_sides=16 | 32)
    return c


CONFIGS = _configs()


@pytest.fixture(scope="module")
def pairs():
    from mrhyde_amd import Mhx
    from oracle import Oracle

    out = {k: (Mhx(p), Oracle(p)) for k, p in CONFIGS.items()}
    yield out
    for g, _ in out.values():
        g.close()


def _state(o, seed):
    return o.setDirichlet(np.random.default_rng(seed).standard_normal(o.n_dof))


def test_default_on_box_q2(pairs):
    g, _ = pairs["c2shape_5x4x3"]
    assert g.get_option("assembly.colouring") == 3


@pytest.mark.parametrize("segment", [2, 4, 6, 12])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_q2box_parity(pairs, name, segment):
    g, o = pairs[name]
    print(f"[q2box] {name} segment={segment} n_dof={g.n_dof}", flush=True)
    g.set_option("assembly.colouring", 3)
    g.set_option("assembly.segment", segment)
    u = _state(o, 31)
    rg = np.full(g.n_dof, 7.0)
    jg = np.full(g.nnz, -3.0)  # stale values must not leak into an overwrite-mode assembly
    g.assemble(u, rg, jg)
    ro, jo = o.assembleJacRes(u)
    rp, _ = o.crs_pattern()
    assert_entries_close(jg, jo, rp, what=f"{name} Jacobian")
    assert_entries_close(rg, ro, None, what=f"{name} residual")
    # repeated assembly: the same bits
    rg2 = np.zeros(g.n_dof)
    jg2 = np.zeros(g.nnz)
    g.assemble(u, rg2, jg2)
    assert np.array_equal(jg, jg2) and np.array_equal(rg, rg2)
    # residual-only pass: the same residual bits
    assert np.array_equal(g.assembleRes(u), rg)
    g.set_option("assembly.segment", 12)


@pytest.mark.parametrize("name", ["c2shape_5x4x3", "modes_7x3x2_all_dbc"])
def test_q2box_accumulate(pairs, name):
    from mrhyde_amd.abi import MHX_BUF_JAC, MHX_BUF_RES, MHX_BUF_U

    g, o = pairs[name]
    g.set_option("assembly.colouring", 3)
    g.set_option("assembly.segment", 4)
    u = _state(o, 5)
    ro, jo = o.assembleJacRes(u)
    g.state_upload(MHX_BUF_U, u)
    g.assemble_resident()
    g.assemble_resident(accumulate=True)
    rp, ci = o.crs_pattern()
    mask, _ = o.dirichlet()
    jo2 = 2 * jo
    diag = np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i) for i in np.where(mask)[0]], dtype=int)
    if len(diag):
        jo2[diag] = 1.0  # dofConstraints replaces the diagonal, it does not add
    assert_entries_close(g.state_download(MHX_BUF_JAC), jo2, rp, what="accumulated Jacobian")
    assert_entries_close(g.state_download(MHX_BUF_RES), 2 * ro, None, what="accumulated residual")
    g.set_option("assembly.segment", 12)


def test_q2box_matches_element_colouring(pairs):
    """The two kernels agree with each other as with the oracle (1e-12)."""
    g, o = pairs["modes_7x3x2_all_dbc"]
    u = _state(o, 9)
    rp, _ = o.crs_pattern()
    g.set_option("assembly.colouring", 0)
    r0, j0 = g.assembleJacRes(u)
    g.set_option("assembly.colouring", 3)
    r3, j3 = g.assembleJacRes(u)
    assert_entries_close(j3, j0, rp, what="pencil vs element colouring Jacobian")
    assert_entries_close(r3, r0, None, what="pencil vs element colouring residual")


def test_q2box_refused_where_it_does_not_apply():
    from mrhyde_amd import Mhx, MhxError
    from mrhyde_amd.abi import Problem

    perturbed = Mhx(Problem(dim=3, order=2, quadrature=4, n=(3, 3, 3), perturb=0.2, perturb_seed=3))
    assert perturbed.get_option("assembly.colouring") == 0
    with pytest.raises(MhxError):
        perturbed.set_option("assembly.colouring", 3)
    perturbed.close()
    q1 = Mhx(Problem(dim=3, order=1, quadrature=2, n=(3, 3, 3)))
    with pytest.raises(MhxError):
        q1.set_option("assembly.colouring", 3)
    q1.close()
