"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/mhx.h, the ctypes layouts match the C layouts, argument
validation fails cleanly, and the seeded generator matches its restatement."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT
from mrhyde_amd import abi

HEADER = os.path.join(ROOT, "include", "mhx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mrhyde_amd import capi

    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(capi.lib, n), f"{n} declared in mhx.h but not exported"
    assert set(names) == set(capi.EXPORTED)
    assert b"gfx950" in capi.lib.mhx_version()


def test_struct_layouts_match_c_header():
    fields = {
        "mhx_problem_desc": abi.ProblemDesc,
        "mhx_krylov_opts": abi.KrylovOpts,
        "mhx_newton_opts": abi.NewtonOpts,
        "mhx_solve_stats": abi.SolveStats,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, py in fields.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_argument_validation_without_device():
    from mrhyde_amd import capi

    h = C.c_void_p()
    bad = abi.Problem(dim=4).desc()
    assert capi.lib.mhx_ctx_create(C.byref(bad), C.byref(h)) == abi.MHX_ERR_ARG
    assert b"dim" in capi.lib.mhx_last_error(None)
    unsupported = abi.Problem(dim=3, order=3, quadrature=6, n=(2, 2, 2)).desc()
    assert capi.lib.mhx_ctx_create(C.byref(unsupported), C.byref(h)) == abi.MHX_ERR_UNSUPPORTED
    assert capi.lib.mhx_assemble(None, None, 0, None, None) == abi.MHX_ERR_ARG
    assert capi.lib.mhx_sizes(None, None, None, None, None) == abi.MHX_ERR_ARG


@pytest.mark.parametrize("kind,a,b", [(0, -1.0, 1.0), (1, 0.0, 0.5), (2, 0.0, 1.0)])
def test_rng_matches_restatement(kind, a, b):
    from mrhyde_amd import capi

    n = 10000
    out = np.empty(n)
    assert capi.lib.mhx_rng_fill(C.c_uint64(12345), n, kind, a, b, out.ctypes.data) == 0
    ref = abi.rng_fill(12345, n, kind, a, b)
    np.testing.assert_allclose(out, ref, rtol=1e-15, atol=0)
    if kind == 0:
        assert out.min() >= a and out.max() < b and abs(out.mean()) < 0.03
    if kind == 2:
        assert abs(out.mean()) < 0.03 and abs(out.std() - 1.0) < 0.03


def test_batch_rejects_bad_boxes():
    """A non-positive box side is refused before any device work; a source
    array of the wrong length is refused by the binding."""
    import numpy as np

    from mrhyde_amd import MhxError, SubgridBatch

    k, g = np.ones(8), np.zeros(27)
    with pytest.raises(MhxError):
        SubgridBatch(np.array([2], dtype=np.int32), k, g, extent=[[1.0, 0.0, 1.0]])
    with pytest.raises(ValueError):
        SubgridBatch(np.array([2], dtype=np.int32), k, g, src=np.ones(7))
