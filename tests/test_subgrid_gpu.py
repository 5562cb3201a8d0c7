"""Batched subgrid solves (one persistent launch) against the CPU oracle.

Each instance is the unit cube in n^3 Q1 hexes with element conductivities
and Dirichlet data on every boundary node; the oracle assembles the same
system through the reference algorithm and solves it with Jacobi-PCG."""
import numpy as np
import pytest

from conftest import spmv_np
from mrhyde_amd.abi import KrylovOpts, MHX_KRYLOV_CG, Problem, subgrid_workload

pytestmark = pytest.mark.gpu


def oracle_instance(n, k, g):
    from oracle import Oracle

    nodal = np.asarray(g, dtype=np.float64)
    o = Oracle(Problem(dim=3, order=1, quadrature=2, n=(n, n, n), k0=k, dbc_values=nodal))
    u0 = o.setDirichlet(np.zeros(o.n_dof))
    res, jac = o.assembleJacRes(u0)
    opts = KrylovOpts()
    opts.solver, opts.tol, opts.max_iters = MHX_KRYLOV_CG, 1e-13, 20000
    du, _ = o.linearSolver(jac, res, opts=opts)
    return o, u0, res, jac, u0 + du


@pytest.fixture(scope="module")
def small_batch():
    rng = np.random.default_rng(12)
    n_axis = np.array([2, 3, 4, 6, 4, 2, 8, 5, 3, 6, 16, 12] * 3, dtype=np.int32)
    k = np.exp(rng.standard_normal(int((n_axis.astype(np.int64) ** 3).sum())))
    g = rng.uniform(-1, 1, int(((n_axis.astype(np.int64) + 1) ** 3).sum()))
    return n_axis, k, g


def check_against_oracle(batch, n_axis, k, g, x, rel, tol_sol=1e-8):
    ko = np.concatenate([[0], np.cumsum(n_axis.astype(np.int64) ** 3)])
    for i, n in enumerate(n_axis):
        print(f"[subgrid] instance {i} n={n} relres={rel[i]:.2e}", flush=True)
        xi = batch.instance(x, i)
        gi = batch.instance(g, i)
        o, u0, res, jac, u_ref = oracle_instance(int(n), k[ko[i]:ko[i + 1]], gi)
        rp, ci = o.crs_pattern()
        r_true = res - spmv_np(rp, ci, jac, xi - u0)
        rel_true = np.linalg.norm(r_true) / np.linalg.norm(res)
        assert rel[i] <= 1e-10 and rel_true <= 2e-10, (i, n, rel[i], rel_true)
        np.testing.assert_allclose(xi, u_ref, rtol=0, atol=tol_sol * np.abs(u_ref).max())
        mask, _ = o.dirichlet()
        np.testing.assert_array_equal(xi[mask == 1], gi[mask == 1])


@pytest.mark.parametrize("wg", [512, 1024])
def test_batch_fp64_matches_oracle(small_batch, wg):
    from mrhyde_amd import SubgridBatch

    n_axis, k, g = small_batch
    b = SubgridBatch(n_axis, k, g, tol=1e-10, wg_size=wg, use_graph=False)
    b.solve()
    x, it, rel, _ = b.results()
    check_against_oracle(b, n_axis, k, g, x, rel)


def test_batch_fp32_refined_matches_oracle(small_batch):
    from mrhyde_amd import SubgridBatch

    n_axis, k, g = small_batch
    b = SubgridBatch(n_axis, k, g, tol=1e-10, precision="fp32-refined", wg_size=512, use_graph=False)
    b.solve()
    x, it, rel, refs = b.results()
    assert refs.min() >= 1 and refs.max() <= 20
    check_against_oracle(b, n_axis, k, g, x, rel)
    b64 = SubgridBatch(n_axis, k, g, tol=1e-10, wg_size=512, use_graph=False)
    b64.solve()
    x64, _, _, _ = b64.results()
    # error of the mixed-precision result against the all-FP64 solve
    assert np.abs(x - x64).max() <= 1e-8 * np.abs(x64).max()


def test_graph_replays_are_identical(small_batch):
    from mrhyde_amd import SubgridBatch

    n_axis, k, g = small_batch
    b = SubgridBatch(n_axis, k, g, tol=1e-10, wg_size=512, use_graph=True)
    eager = SubgridBatch(n_axis, k, g, tol=1e-10, wg_size=512, use_graph=False)
    eager.solve()
    xe, ite, _, _ = eager.results()
    for _ in range(3):
        b.solve()
        x, it, rel, _ = b.results()
        np.testing.assert_array_equal(x, xe)
        np.testing.assert_array_equal(it, ite)


def test_full_batch_properties():
    """The 4096-instance workload at full size: every instance reaches the
    tolerance, Dirichlet values are kept, and the solution obeys the discrete
    maximum principle (the unit-cube Q1 stiffness is an M-matrix, no source)."""
    from mrhyde_amd import SubgridBatch

    n_axis, k, g = subgrid_workload()
    b = SubgridBatch(n_axis, k, g, tol=1e-10, use_graph=True)
    ms = b.solve()
    x, it, rel, _ = b.results()
    print(f"[subgrid] 4096 instances in {ms:.2f} ms, wg={b.wg_size}, iters {it.min()}..{it.max()}", flush=True)
    assert rel.max() <= 1e-10
    for i in range(0, len(n_axis), 97):
        n = int(n_axis[i]) + 1
        xi = b.instance(x, i).reshape(n, n, n)
        gi = b.instance(g, i).reshape(n, n, n)
        bnd = np.ones((n, n, n), dtype=bool)
        bnd[1:-1, 1:-1, 1:-1] = False
        np.testing.assert_array_equal(xi[bnd], gi[bnd])
        assert xi.min() >= gi[bnd].min() - 1e-9 and xi.max() <= gi[bnd].max() + 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32-refined"])
def test_box_and_source_batch_matches_oracle(precision):
    """Instances on boxes of their own side lengths with element sources:
    the box stiffness (cx S(x)M(x)M + ...) and the source load f_e |e| / 8
    against the oracle's quadrature assembly of the same box and source."""
    from oracle import Oracle
    from mrhyde_amd import SubgridBatch

    rng = np.random.default_rng(21)
    n_axis = np.array([2, 3, 5, 8, 4, 6, 16, 3], dtype=np.int32)
    ext = rng.uniform(0.3, 2.5, (len(n_axis), 3))
    ne = n_axis.astype(np.int64) ** 3
    k = np.exp(rng.standard_normal(int(ne.sum())))
    f = rng.uniform(-5.0, 5.0, int(ne.sum()))
    g = rng.uniform(-1, 1, int(((n_axis.astype(np.int64) + 1) ** 3).sum()))
    b = SubgridBatch(n_axis, k, g, tol=1e-10, precision=precision, use_graph=False, extent=ext, src=f)
    b.solve()
    x, it, rel, _ = b.results()
    ko = np.concatenate([[0], np.cumsum(ne)])
    for i, n in enumerate(n_axis):
        print(f"[subgrid box] instance {i} n={n} L={ext[i]} relres={rel[i]:.2e}", flush=True)
        xi, gi = b.instance(x, i), b.instance(g, i)
        o = Oracle(Problem(dim=3, order=1, quadrature=2, n=(int(n),) * 3, hi=tuple(ext[i]), k0=k[ko[i]:ko[i + 1]],
                           src_elem=f[ko[i]:ko[i + 1]], dbc_values=np.asarray(gi)))
        u0 = o.setDirichlet(np.zeros(o.n_dof))
        res, jac = o.assembleJacRes(u0)
        opts = KrylovOpts()
        opts.solver, opts.tol, opts.max_iters = MHX_KRYLOV_CG, 1e-13, 20000
        du, _ = o.linearSolver(jac, res, opts=opts)
        u_ref = u0 + du
        rp, ci = o.crs_pattern()
        r_true = res - spmv_np(rp, ci, jac, xi - u0)
        assert rel[i] <= 1e-10 and np.linalg.norm(r_true) <= 2e-10 * np.linalg.norm(res), (i, rel[i])
        np.testing.assert_allclose(xi, u_ref, rtol=0, atol=1e-8 * np.abs(u_ref).max())
