"""The declared arithmetic and the basic fp32 ops give the same bits on gfx950 as on the
host (correctly rounded div/sqrt, no contraction): the precondition of per-pixel bit parity."""
import ctypes as C

import numpy as np
import pytest

from pathtracerglsl_amd import _lib

pytestmark = pytest.mark.gpu


def gpu(fn, x, y):
    lib = _lib.load()
    lib.pt_debug_math.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    out = np.empty_like(x)
    _lib.check(lib.pt_debug_math(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, len(x)))
    return out


def cpu(oracle, fn, x, y):
    out = np.empty_like(x)
    f = oracle.or_math_eval if fn < 20 else oracle.or_basic_eval
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    f(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, len(x))
    return out


CASES = [(0, -30, 30), (1, -30, 30), (2, -80, 80), (3, 1e-6, 1e4), (4, -5, 5), (5, -1.5, 1.5), (6, 0, 1), (6, -1, 0),
         (7, -40, 40), (8, -1, 1), (9, -6, 6), (10, -0.99, 0.99), (11, -1, 1), (20, -5, 5), (21, 0, 100), (22, 1e-3, 100),
         (23, 0, 100), (24, 0.01, 10), (25, -3, 3), (26, -3, 3)]


@pytest.mark.parametrize("fn,lo,hi", CASES)
def test_device_bits_equal_host_bits(oracle, fn, lo, hi):
    rng = np.random.default_rng(fn)
    x = rng.uniform(lo, hi, 1 << 16).astype(np.float32)
    y = rng.uniform(lo, hi, 1 << 16).astype(np.float32)
    if fn == 6:
        y[:] = 5.0
    g, c = gpu(fn, x, y), cpu(oracle, fn, x, y)
    diff = np.nonzero(g.view(np.uint32) != c.view(np.uint32))[0]
    print(f"fn {fn}: {len(diff)} differing of {len(x)}", flush=True)
    assert len(diff) == 0, (x[diff[:5]], y[diff[:5]], g[diff[:5]], c[diff[:5]])
