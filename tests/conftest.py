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
