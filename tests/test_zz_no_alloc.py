"""m4g_dot_noalloc, the entry point that allocates nothing, against m4g_dot
through the C ABI: two calls in a row (the second on a smaller input, so the
scratch array still holds the first call's partials) must return m4g_dot's
result bit for bit, and the oracle's to 1e-12 relative."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def c_dot(lib, name, x, y):
    out = ctypes.c_double(np.nan)
    st = getattr(lib, name)(ctypes.c_void_p(x.h.value), ctypes.c_void_p(y.h.value), ctypes.byref(out))
    assert st == 0, (name, st)
    return out.value


def test_dot_noalloc_matches_dot(g):
    lib = ctypes.CDLL(os.path.join(os.path.dirname(g.__file__), "lib", "libmtl4gpu.so"))
    for name in ("m4g_dot", "m4g_dot_noalloc"):
        getattr(lib, name).restype = ctypes.c_int
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    assert lib.m4g_set_device(0) == 0
    # 3,000,017 entries use all 1024 partials; 777 entries use one, so the
    # second call reads only what it wrote itself
    for n, seed in ((3_000_017, 11), (777, 12)):
        print(f"dot_noalloc n={n}", flush=True)
        xh, yh = g.rng_uniform(seed, 1, n), g.rng_uniform(seed, 2, n)
        x, y = g.Vector(data=xh), g.Vector(data=yh)
        got = c_dot(lib, "m4g_dot_noalloc", x, y)
        want = c_dot(lib, "m4g_dot", x, y)
        assert got == want, (n, got, want)
        ref = o.dot(xh, yh)
        assert abs(got - ref) <= 1e-12 * np.sum(np.abs(xh * yh))
        assert g.dot_noalloc(x, y) == want


def test_dot_noalloc_rejects_bad_arguments(g):
    x, y = g.Vector(data=np.ones(10)), g.Vector(data=np.ones(11))
    with pytest.raises(g.M4GError) as e:
        g.dot_noalloc(x, y)
    assert e.value.status == g.ERR_SIZE
    xf = g.Vector(10, dtype=np.float32)
    with pytest.raises(g.M4GError):
        g.dot_noalloc(xf, xf)
