"""ghx_rng_gauss_noalloc (no device allocation, staging in a file-scope device
array) against ghx_rng_gauss through the C ABI: equal draws, bit for bit, on
two calls in a row, the second with a smaller input; sizes beyond the
array's 2^20 draws are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_DRAWS = 1 << 20


def draws(lib, fn, seed, key0, n):
    out = np.full(n, np.nan)
    rc = getattr(lib, fn)(seed, key0, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_equals_allocating_entry_point():
    from gauge_higgs_amd import _lib
    lib = _lib.load()
    for seed, key0, n in [(12345, 0, MAX_DRAWS), (7, 10 ** 9 + 3, 1000)]:
        rc, a = draws(lib, "ghx_rng_gauss_noalloc", seed, key0, n)
        assert rc == 0
        rc, b = draws(lib, "ghx_rng_gauss", seed, key0, n)
        assert rc == 0
        assert np.array_equal(a, b)
        print(f"noalloc: seed {seed} key0 {key0} n {n}: equal")
    # the second, smaller call left the rest of the array as the first wrote it;
    # only the range asked for is returned
    rc, _ = draws(lib, "ghx_rng_gauss_noalloc", 1, 0, MAX_DRAWS + 1)
    assert rc == -1
    assert lib.ghx_rng_gauss_noalloc(1, 0, -1, None) == -1
    assert lib.ghx_rng_gauss_noalloc(1, 0, 0, None) == 0
