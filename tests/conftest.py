import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the libmhx.so kernels")
    config.addinivalue_line("markers", "slow: full-size configuration")


def pytest_collection_finish(session):
    """PyTorch ships its own HIP runtime beside the system one that libmhx.so
    links.  When libmhx.so brings the device up first, torch reports no device
    for the rest of the process, so the GPU tests that hand torch tensors to
    the engine need torch's runtime started before any other test runs."""
    if not any(item.get_closest_marker("gpu") for item in session.items):
        return
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def spmv_np(rowptr, colidx, vals, x):
    """Reference CSR product in numpy (checker only)."""
    y = np.zeros(len(rowptr) - 1)
    np.add.at(y, np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr)), vals * x[colidx])
    return y


def to_scipy(rowptr, colidx, vals):
    import scipy.sparse as sp

    n = len(rowptr) - 1
    return sp.csr_matrix((vals, colidx, rowptr), shape=(n, n))


def assert_entries_close(a, b, rowptr, rel=1e-12, floor=1e-14, what="entries"):
    """|a - b| <= rel*|b| + floor*rowmax(|b|) entry-wise (rowmax: largest |b| in the
    entry's CRS row, or the vector's largest entry when rowptr is None)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if rowptr is None:
        scale = np.full_like(b, np.abs(b).max())
    else:
        rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
        rmax = np.zeros(len(rowptr) - 1)
        np.maximum.at(rmax, rows, np.abs(b))
        scale = rmax[rows]
    err = np.abs(a - b)
    bound = rel * np.abs(b) + floor * scale
    bad = err > bound
    if bad.any():
        i = int(np.argmax(err / np.maximum(bound, 1e-300)))
        raise AssertionError(f"{what}: {int(bad.sum())} entries out of tolerance; worst at {i}: "
                             f"{a[i]!r} vs {b[i]!r} (bound {bound[i]:.3e})")


def field_l2_errors(oracle, u, exact):
    """L2 error of every field at the volume cubature points, as the reference's
    PostprocessManager::computeError does (sum over elements and points of
    w |det J| (u_h - u_true)^2, square root at print); exact[f](x) -> values.
    Uses the oracle's physical basis and weights (checker only)."""
    nf = oracle.problem.n_field
    L = oracle.lids()
    err = np.zeros(nf)
    for e in range(oracle.n_elem):
        phi, _, wts, x = oracle.elem_basis(e)
        for f in range(nf):
            uh = u[L[e, f * oracle.nb:(f + 1) * oracle.nb]] @ phi
            err[f] += np.sum(wts * (uh - exact[f](x)) ** 2)
    return np.sqrt(err)


def ns_channel_problem(deck, nz=0, depth=0.0):
    """The channel deck (tests/golden/ns_gold.json); nz > 0 extrudes it in z
    with uz = 0 on the z faces (ux, uy free there)."""
    from mrhyde_amd.abi import MHX_PHYS_NAVIER_STOKES, Problem

    n, lo, hi = list(deck["n"]), list(deck["lo"]), list(deck["hi"])
    dim = deck["dim"]
    comp = (0, 0, 0)
    if nz:
        dim, n, lo, hi = 3, n + [nz], lo + [0.0], hi + [depth]
        comp = (0, 0, 16 | 32)
    return Problem(dim=dim, order=deck["order"], quadrature=deck["quadrature"], n=n, lo=lo, hi=hi,
                   physics=MHX_PHYS_NAVIER_STOKES, dbc_sides=deck["vel_dbc_sides"], rho=deck["rho"],
                   mu=deck["mu"], pspg=deck["pspg"], supg=deck["supg"], force=deck["force"], comp_sides=comp)


def thermal_gold_problem(g):
    """A thermal gold case of tests/golden/thermal_gold.json as a Problem, with
    its true solution as (amplitude, frequencies) terms."""
    from mrhyde_amd.abi import Problem

    PI = np.pi
    modes = [(a * PI * PI, [w * PI for w in f]) for a, f in g["source"]]
    flux = [([a * PI for a in amp], [w * PI for w in fr], [ph * PI for ph in phs])
            for amp, fr, phs in g.get("neumann_flux", [])]
    p = Problem(dim=g["dim"], order=g["order"], quadrature=g["quadrature"], n=tuple(g["n"]), modes=modes,
                workset_size=g["workset"], dbc_sides=g.get("dbc_sides", 63),
                neumann_sides=g.get("neumann_sides", 0), neumann_flux=flux)
    true = [(a, [w * PI for w in f]) for a, f in g["true"]]
    return p, true, g
