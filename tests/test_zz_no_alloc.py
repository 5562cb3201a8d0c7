"""mhx_linear_solve_noalloc against mhx_linear_solve through the C ABI.

The allocation-free entry point runs Jacobi-PCG / CG with its Krylov vectors
in the library's static per-GPU scratch array; with the same options it must
give the same bits as the allocating entry point, and meet the linear-solve
bar (relative residual <= 1e-10 against the oracle's Jacobian).  Two calls in
a row, the second on a smaller problem, so the second starts from the first
call's leftovers in the array.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import spmv_np

pytestmark = pytest.mark.gpu


def _solve(lib, fn, g, jac, r, x, opts):
    from mrhyde_amd import capi

    st = capi.SolveStats()
    rc = getattr(lib, fn)(g._h, C.c_void_p(jac.data_ptr()), C.c_void_p(r.data_ptr()), C.c_void_p(x.data_ptr()),
                          C.byref(opts), C.byref(st))
    return rc, st


def _case(problem, precond):
    import torch

    from mrhyde_amd import Mhx, capi, krylov_opts
    from oracle import Oracle

    lib = capi.lib
    g, o = Mhx(problem), Oracle(problem)
    u = o.setDirichlet(np.zeros(o.n_dof))
    r_h, j_h = g.assembleJacRes(u)
    r_o, j_o = o.assembleJacRes(u)
    dev = torch.device("cuda:0")
    jac = torch.from_numpy(j_h).to(dev)
    r = torch.from_numpy(r_h).to(dev)
    xa = torch.zeros(g.n_dof, dtype=torch.float64, device=dev)
    xb = torch.zeros(g.n_dof, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    opts = krylov_opts(solver="cg", precond=precond, tol=1e-10, max_iters=20000, device=True)
    print(f"[no_alloc] n_dof={g.n_dof} precond={precond}", flush=True)
    rca, sta = _solve(lib, "mhx_linear_solve", g, jac, r, xa, opts)
    rcb, stb = _solve(lib, "mhx_linear_solve_noalloc", g, jac, r, xb, opts)
    assert rca == 0 and rcb == 0, (rca, rcb, lib.mhx_last_error(g._h))
    xa_h, xb_h = xa.cpu().numpy(), xb.cpu().numpy()
    # same kernels in the same order: the same bits
    assert np.array_equal(xa_h, xb_h)
    assert sta.iters == stb.iters and sta.rel_res == stb.rel_res
    rp, ci = o.crs_pattern()
    rel = np.linalg.norm(r_o - spmv_np(rp, ci, j_o, xb_h)) / np.linalg.norm(r_o)
    assert stb.rel_res <= 1e-10 and rel <= 1.5e-10, (stb.rel_res, rel)
    return g, jac, r, xb, opts


def test_noalloc_matches_allocating_solve_twice():
    from mrhyde_amd.workloads import poisson_2d
    from mrhyde_amd.abi import Problem

    # a 3-D Q1 problem with a lognormal field first, then the smaller 2-D
    # 64 x 64 Poisson problem: the second call finds the first call's
    # vectors and scalars in the array
    k0 = np.exp(0.5 * np.random.default_rng(5).standard_normal(20 ** 3))
    big = Problem(dim=3, order=1, quadrature=2, n=(20, 20, 20), k0=k0,
                  modes=[(1.0, (np.pi, np.pi, np.pi))])
    g1, *_ = _case(big, "jacobi")
    g2, *_ = _case(poisson_2d(64), "jacobi")
    g3, *_ = _case(poisson_2d(24), "none")
    for g in (g1, g2, g3):
        g.close()


def test_noalloc_rejects_what_it_cannot_do():
    import torch

    from mrhyde_amd import Mhx, capi, krylov_opts
    from mrhyde_amd.workloads import poisson_2d

    lib = capi.lib
    g = Mhx(poisson_2d(16))
    u = np.zeros(g.n_dof)
    r_h, j_h = g.assembleJacRes(u)
    dev = torch.device("cuda:0")
    jac = torch.from_numpy(j_h).to(dev)
    r = torch.from_numpy(r_h).to(dev)
    x = torch.zeros(g.n_dof, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    st = capi.SolveStats()
    # host pointers
    xh = np.zeros(g.n_dof)
    o = krylov_opts(solver="cg", precond="jacobi", tol=1e-10, max_iters=1000, device=False)
    rc = lib.mhx_linear_solve_noalloc(g._h, j_h.ctypes.data_as(C.c_void_p), r_h.ctypes.data_as(C.c_void_p),
                                      xh.ctypes.data_as(C.c_void_p), C.byref(o), C.byref(st))
    assert rc == capi.MHX_ERR_ARG
    # GMRES, AMG and the FP32 inner solve are not built allocation-free
    for kw in (dict(solver="gmres", precond="jacobi"), dict(solver="cg", precond="amg"),
               dict(solver="cg", precond="jacobi-fp32")):
        o = krylov_opts(tol=1e-10, max_iters=1000, device=True, **kw)
        rc, _ = _solve(lib, "mhx_linear_solve_noalloc", g, jac, r, x, o)
        assert rc == capi.MHX_ERR_UNSUPPORTED, kw
    g.close()
    # above the stated maximum size
    big = Mhx(poisson_2d(1024))  # 1,050,625 DOFs > 1,048,576
    assert big.n_dof > 1048576
    o = krylov_opts(solver="cg", precond="jacobi", tol=1e-10, max_iters=10, device=True)
    rc, _ = _solve(lib, "mhx_linear_solve_noalloc", big, r, r, x, o)
    assert rc == capi.MHX_ERR_ARG
    assert b"MHX_NOALLOC_MAX_DOF" in lib.mhx_last_error(big._h)
    big.close()
