"""tasep_moments_noalloc (the allocation-free variant of tasep_moments)
against tasep_moments, through the C ABI: two calls in a row, the second on
a smaller lattice, and the argument checks that keep it from launching."""
import ctypes as C

import numpy as np
import pytest

import tasep_amd as T

pytestmark = pytest.mark.gpu


def _noalloc(ptr, n, prec, stream):
    L = T.load_library()
    y = np.zeros(n)
    cov = np.zeros(n * n)
    mul = np.zeros(n * n)
    m = np.zeros(1)
    rc = L.tasep_moments_noalloc(C.c_void_p(ptr), n, prec, y.ctypes.data, cov.ctypes.data,
                                 mul.ctypes.data, m.ctypes.data, C.c_void_p(stream))
    return rc, y, cov.reshape(n, n), mul.reshape(n, n), float(m[0])


def test_moments_noalloc_matches_moments_twice():
    """Second call smaller than the first: nothing the first call left in
    the shared scratch may reach the second's results."""
    for n, prec, seed in ((22, "f64", 1), (13, "f32", 2)):
        rng = np.random.default_rng(seed)
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_rates(rng.uniform(0.2, 1.8, n + 1))
            e.set_x0_product(rng.uniform(0.15, 0.85, n))
            e.run(0.0, 0.01, 25, patience=-1)
            y0, cov0, mul0, m0 = e.moments()
            print(f"n={n} {prec}: tasep_moments_noalloc vs tasep_moments", flush=True)
            rc, y, cov, mul, m = _noalloc(e.x_device_ptr, n, e.precision, e.stream)
        assert rc == 0, T.load_library().tasep_strerror(rc)
        assert np.max(np.abs(y - y0)) <= 1e-13
        assert np.max(np.abs(cov - cov0)) <= 1e-13
        assert np.max(np.abs(mul - mul0)) <= 1e-13
        assert abs(m - m0) <= 1e-13


def test_moments_noalloc_rejects_bad_arguments():
    with T.Engine(10, "f64", "rk4", "mf") as e:
        e.set_x0_uniform()
        ptr = e.x_device_ptr
        assert _noalloc(ptr, 1, T.F64, e.stream)[0] == -1          # TASEP_ERR_ARG
        assert _noalloc(ptr, 35, T.F64, e.stream)[0] < 0
        assert _noalloc(ptr, 10, 7, e.stream)[0] < 0
        assert _noalloc(0, 10, T.F64, e.stream)[0] < 0
        host = np.zeros(1 << 10)
        assert _noalloc(host.ctypes.data, 10, T.F64, e.stream)[0] < 0
        rc, y, _, _, m = _noalloc(ptr, 10, T.F64, e.stream)   # still usable afterwards
        assert rc == 0 and np.allclose(y, 0.5, rtol=0, atol=1e-15) and abs(m - 1.0) < 1e-15
