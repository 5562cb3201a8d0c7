"""Serving loop (mhx_server_*): continuous batching of independent thermal
solves of mixed size.  Each result must be the same alone and in the mix,
solve to the requested tolerance, and the server must respect its device-
memory budget and reuse resident contexts."""
import numpy as np
import pytest

from mrhyde_amd.abi import Problem

pytestmark = pytest.mark.gpu
PI = np.pi


def mixed_requests(seed, count):
    """2-D Q1 and 3-D Q1/Q2 thermal problems with lognormal k (sigma 0.5) and
    a 1-4 mode sinusoidal source; a few mesh shapes repeat with new data."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 5 == 4:  # same mesh as an earlier request, new coefficients
            prev = out[int(rng.integers(0, len(out)))]
            dim, order, n = prev.dim, prev.order, prev.n
        elif rng.random() < 0.5:
            dim, order, n = 2, 1, (int(rng.integers(6, 40)),) * 2
        else:
            dim, order, n = 3, int(rng.integers(1, 3)), (int(rng.integers(3, 10)),) * 3
        ne = int(np.prod(n))
        modes = [(float(rng.uniform(0, 1)), [PI * int(rng.integers(1, 5)) for _ in range(dim)])
                 for _ in range(int(rng.integers(1, 5)))]
        out.append(Problem(dim=dim, order=order, quadrature=2 * order, n=n,
                           k0=np.exp(0.5 * rng.standard_normal(ne)), modes=modes))
    return out


def test_results_same_alone_and_in_mix():
    from mrhyde_amd import Server
    from oracle import Oracle

    probs = mixed_requests(1, 30)
    alone = []
    s1 = Server(max_active=1, max_cached=0)
    for p in probs:
        alone.append(s1.poll(s1.submit(p))[0])
    s1.close()
    s8 = Server(max_active=6, max_cached=8)
    tickets = [s8.submit(p) for p in probs]
    mix = [s8.poll(t) for t in tickets]
    info = s8.info()
    print(f"[serve] mix info {info}", flush=True)
    assert info["reused"] >= 1
    for i, (p, xa, (xm, st, lat)) in enumerate(zip(probs, alone, mix)):
        # the stated tolerance: equal to 1e-12 of the solution's size (bit-identical in practice)
        np.testing.assert_allclose(xm, xa, rtol=0, atol=1e-12 * np.abs(xa).max(), err_msg=f"request {i}")
        assert lat > 0
        if i % 6 == 0:  # solved: the oracle's residual at the result is at the Newton tolerance
            o = Oracle(p)
            r0 = np.abs(o.assembleRes(o.setDirichlet(np.zeros(o.n_dof)))).max()
            assert np.abs(o.assembleRes(xm)).max() <= 1e-8 * r0, i
    s8.close()


def test_admission_by_footprint():
    from mrhyde_amd import MhxError, Server

    big = Problem(dim=3, order=2, quadrature=4, n=(12, 12, 12))
    small = [Problem(dim=2, order=1, quadrature=2, n=(20 + i, 20 + i), modes=[(1.0, (PI, PI))]) for i in range(8)]
    _, _, b_big = Server.footprint(big)
    _, _, b_small = Server.footprint(small[-1])
    budget = 3 * b_small  # room for about three of the small requests at a time
    assert b_big > budget
    s = Server(max_active=6, max_cached=2, hbm_budget=budget)
    with pytest.raises(MhxError) as ei:
        s.submit(big)
    assert ei.value.code == -6
    tickets = [s.submit(p) for p in small]
    for t in tickets:
        x, st, lat = s.poll(t)
        assert st["rel_res"] <= 1e-10
    info = s.info()
    assert info["done"] == len(small) and info["bytes_held"] <= budget
    s.close()


def test_reused_context_matches_fresh_context():
    from mrhyde_amd import Mhx, Server

    rng = np.random.default_rng(7)
    shape = dict(dim=3, order=2, quadrature=4, n=(5, 4, 3))
    probs = [Problem(**shape, k0=np.exp(0.5 * rng.standard_normal(60)), src_elem=rng.uniform(0, 1, 60))
             for _ in range(3)]
    s = Server(max_active=1, max_cached=4)
    got = [s.poll(s.submit(p))[0] for p in probs]
    assert s.info()["reused"] == 2 and s.info()["created"] == 1
    s.close()
    for p, x in zip(probs, got):
        g = Mhx(p)
        u = g.setDirichlet(np.zeros(g.n_dof))
        u, st = g.nonlinearSolver(u, max_iters=2, nl_tol=1e-8, solver="cg", tol=1e-10, linear_max_iters=20000)
        np.testing.assert_allclose(x, u, rtol=0, atol=1e-12 * np.abs(u).max())
