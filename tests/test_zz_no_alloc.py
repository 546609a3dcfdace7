"""sc_poisson_solve_noalloc (the vector-potential solve on a static device
scratch array, for callers that cannot allocate) against sc_poisson_solve,
through the C ABI: two calls in a row, the second on a smaller grid, each
equal to the allocating entry point's result; sizes beyond the maximum are
refused with a clear error before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_DP = ctypes.POINTER(ctypes.c_double)


def _noalloc(L, j, h, ng, enforce=1):
    n = j.shape[-1]
    j = np.ascontiguousarray(j, dtype=np.float64)
    out = np.full((2, n, n), np.nan)
    rc = L.sc_poisson_solve_noalloc(n, ctypes.c_double(h), ng, enforce,
                                    j.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    return rc, out


def test_noalloc_matches_allocating_solve(product):
    L = product.lib()
    L.sc_poisson_solve_noalloc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, _DP, _DP]
    rng = np.random.default_rng(21)
    # the largest grid first, then a smaller one: the second call must not
    # see anything the first left in the shared scratch
    for n, h, ng in ((512, 0.05, 150000), (300, 0.05, 50000), (37, 0.2, 900)):
        print("noalloc case n=%d h=%g" % (n, h), flush=True)
        j = rng.normal(size=(2, n, n))
        ref = product.poisson_solve(j, h, ng)
        rc, out = _noalloc(L, j, h, ng)
        assert rc == 0, L.sc_last_error().decode()
        err = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
        print("n=%d: max rel difference to sc_poisson_solve %.2e" % (n, err), flush=True)
        assert err <= 1e-13


def test_noalloc_refuses_grids_beyond_the_scratch(product):
    L = product.lib()
    L.sc_poisson_solve_noalloc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, _DP, _DP]
    j = np.zeros((2, 513, 513))
    rc, out = _noalloc(L, j, 0.05, 1000)
    assert rc == 1  # SC_ERR_ARGUMENT
    assert "513" in L.sc_last_error().decode()
    assert np.all(np.isnan(out))
