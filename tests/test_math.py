"""The declared fp32 arithmetic (include/ptmi_math.h) against correctly rounded results.
A dropped Taylor term, a wrong Cody-Waite constant or a swapped quadrant shows up as a
large ulp error here."""
import numpy as np
import pytest


def ulp_err(got, ref64):
    got = got.astype(np.float64)
    ref32 = ref64.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).smallest_subnormal))
    return np.abs(got - ref64) / spacing


def run(oracle, fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.empty_like(x)
    oracle.or_math_eval(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, len(x))
    return out


@pytest.mark.parametrize("fn,ref,lo,hi,bound", [
    (0, np.sin, -20.0, 20.0, 2.0),
    (1, np.cos, -20.0, 20.0, 2.0),
    (2, np.exp, -80.0, 80.0, 2.0),
    (7, np.arctan, -50.0, 50.0, 2.0),
    (8, np.arccos, -1.0, 1.0, 3.0),
])
def test_unary_ulp(oracle, fn, ref, lo, hi, bound):
    x = np.random.default_rng(fn).uniform(lo, hi, 200000).astype(np.float32)
    got = run(oracle, fn, x)
    err = ulp_err(got, ref(x.astype(np.float64)))
    if fn in (0, 1):
        # absolute error near the zeros of sin/cos (relative ulp is unbounded there)
        ok = (err <= bound) | (np.abs(got.astype(np.float64) - ref(x.astype(np.float64))) < 2e-7 * max(1.0, abs(hi)))
        assert ok.all(), float(err[~ok].max())
    else:
        assert err.max() <= bound, float(err.max())


def test_log_ulp(oracle):
    x = np.concatenate([np.random.default_rng(1).uniform(1e-6, 1e6, 200000),
                        np.random.default_rng(2).uniform(0.5, 2.0, 100000)]).astype(np.float32)
    got = run(oracle, 3, x)
    ref = np.log(x.astype(np.float64))
    err = ulp_err(got, ref)
    ok = (err <= 2.0) | (np.abs(got - ref) < 1e-7)
    assert ok.all()


def test_atan2_quadrants_and_special(oracle):
    rng = np.random.default_rng(4)
    y = rng.normal(size=100000).astype(np.float32)
    x = rng.normal(size=100000).astype(np.float32)
    got = run(oracle, 4, x, y)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.abs(got - ref).max() < 1e-6
    sp = run(oracle, 4, np.array([0, -1, 1, 0, 0], np.float32), np.array([0, 0, 0, 1, -1], np.float32))
    np.testing.assert_allclose(sp, [0.0, np.pi, 0.0, np.pi / 2, -np.pi / 2], atol=1e-7)


def test_tan_pow_tanh_atanh(oracle):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1.4, 1.4, 50000).astype(np.float32)
    assert (ulp_err(run(oracle, 5, x), np.tan(x.astype(np.float64))) <= 4).all()
    b = rng.uniform(0.0, 1.0, 50000).astype(np.float32)
    got = run(oracle, 6, b, np.full_like(b, 5.0))
    ref = b.astype(np.float64) ** 5
    # exp(y log x) carries the rounding of y*log(x): |y ln x| * 2^-23 relative, plus a few ulp
    bound = (np.abs(5.0 * np.log(np.maximum(b.astype(np.float64), 1e-30))) + 4.0) * 1.2e-7
    assert (np.abs(got - ref) <= bound * ref + 1e-37).all()
    # negative base: integer exponents give the real power, others NaN (GLSL: undefined)
    neg = run(oracle, 6, np.array([-1.1920929e-07, -0.5, -2.0, -0.5], np.float32), np.array([5.0, 2.0, 3.0, 0.5], np.float32))
    tiny = float(np.float32(1.1920929e-07)) ** 5
    assert neg[0] < 0.0 and abs(neg[0] + tiny) <= (abs(5.0 * np.log(1.1920929e-07)) + 4.0) * 1.2e-7 * tiny
    np.testing.assert_allclose(neg[1:3], [0.25, -8.0], rtol=2e-6)
    assert np.isnan(neg[3])
    t = rng.uniform(-5, 5, 50000).astype(np.float32)
    assert np.abs(run(oracle, 9, t) - np.tanh(t.astype(np.float64))).max() < 3e-7
    a = rng.uniform(-0.99, 0.99, 50000).astype(np.float32)
    assert np.abs(run(oracle, 10, a) - np.arctanh(a.astype(np.float64))).max() < 2e-6
    s = np.concatenate([rng.uniform(-1.0, 1.0, 50000), [-1.0, -0.5, 0.0, 0.5, 1.0]]).astype(np.float32)
    got = run(oracle, 11, s)
    assert np.abs(got - np.arcsin(s.astype(np.float64))).max() < 2e-6
    np.testing.assert_allclose(got[-5:], [-np.pi / 2, -np.pi / 6, 0.0, np.pi / 6, np.pi / 2], atol=3e-7)
