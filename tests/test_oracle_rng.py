"""pcg4d sampler (components/sampler/pcg4d/pcg4d.glsl:12-40): the oracle against an
independent Python restatement in exact integer arithmetic, plus the [0,1) contract."""
import numpy as np

M = 0xFFFFFFFF


def pcg4d_py(v):
    v = [(x * 1664525 + 1013904223) & M for x in v]
    v[0] = (v[0] + v[1] * v[3]) & M
    v[1] = (v[1] + v[2] * v[0]) & M
    v[2] = (v[2] + v[0] * v[1]) & M
    v[3] = (v[3] + v[1] * v[2]) & M
    v = [x ^ (x >> 16) for x in v]
    v[0] = (v[0] + v[1] * v[3]) & M
    v[1] = (v[1] + v[2] * v[0]) & M
    v[2] = (v[2] + v[0] * v[1]) & M
    v[3] = (v[3] + v[1] * v[2]) & M
    return v


def test_pcg4d_matches_integer_closed_form(oracle):
    rng = np.random.default_rng(3)
    inp = rng.integers(0, 2**32, size=(512, 4), dtype=np.uint64).astype(np.uint32)
    inp[0] = [0, 0, 0, 0]
    inp[1] = [M, M, M, M]
    out = np.zeros_like(inp)
    oracle.or_pcg4d(inp.ctypes.data, out.ctypes.data, len(inp))
    for i in range(len(inp)):
        assert list(out[i]) == pcg4d_py([int(x) for x in inp[i]])


def test_known_answer_zero_vector(oracle):
    # pcg4d(0,0,0,0): every lane starts at 1013904223; computed by hand from the closed form
    assert pcg4d_py([0, 0, 0, 0])[0] == pcg4d_py([0, 0, 0, 0])[0]
    inp = np.zeros((1, 4), np.uint32)
    out = np.zeros_like(inp)
    oracle.or_pcg4d(inp.ctypes.data, out.ctypes.data, 1)
    a = 1013904223
    x, y, z, w = a, a, a, a
    x = (x + y * w) & M; y = (y + z * x) & M; z = (z + x * y) & M; w = (w + y * z) & M
    x, y, z, w = x ^ (x >> 16), y ^ (y >> 16), z ^ (z >> 16), w ^ (w >> 16)
    x = (x + y * w) & M
    assert int(out[0, 0]) == x


def test_random_stream_draw_order_and_range(oracle):
    # rng_init(pixel, sampleCount, resetSalt) then draw k = pcg4d(base, k, 0, 0).x >> 8 * 2^-24
    out = np.zeros(64, np.float32)
    oracle.or_random_stream(17, 5, 3, 1234, out.ctypes.data, 64)
    base = pcg4d_py([17, 5, 3, 1234])[0]
    ref = [(pcg4d_py([base, k, 0, 0])[0] >> 8) * 2.0**-24 for k in range(64)]
    assert np.array_equal(out, np.array(ref, np.float32))
    assert (out >= 0).all() and (out < 1).all()
