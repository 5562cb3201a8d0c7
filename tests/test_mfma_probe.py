"""The dense matrix-core form of the Q2 element contraction (the measured
alternative to the sum-factorised column stage): both variants reproduce
G diag(d) G^T, checked on the host inside the entry point, on a launch with
a ragged last workgroup and on a single element."""
import ctypes as C

import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n_elem", [1, 1001])
def test_mfma_contraction_matches_host(variant, n_elem):
    from mrhyde_amd import capi

    ms, err = C.c_double(), C.c_double()
    assert capi.lib.mhx_mfma_probe(n_elem, 2, variant, C.byref(ms), C.byref(err)) == 0
    assert err.value < 1e-13
    assert ms.value > 0


def test_mfma_probe_rejects_bad_arguments():
    from mrhyde_amd import capi

    ms, err = C.c_double(), C.c_double()
    assert capi.lib.mhx_mfma_probe(0, 1, 0, C.byref(ms), C.byref(err)) == -1
    assert capi.lib.mhx_mfma_probe(10, 1, 2, C.byref(ms), C.byref(err)) == -1
    assert capi.lib.mhx_mfma_probe(10, 0, 0, C.byref(ms), C.byref(err)) == -1
