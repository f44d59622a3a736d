"""The environment sampler's guided CDF search on the GPU.

The device sampler narrows each of its two searches (the marginal over rows, then the row's
conditional CDF) with a cut-point guide (Chen and Asau 1974): for bucket b of K equal parts of
[0, 1), the index the full search returns at b / K. The reference sampler
(components/env/sampler_cdf.glsl:61-80) searches the whole CDF, and so does the oracle; the
guided search must return the same index for every u. pt_debug_env_sample's directions and pdfs
are compared with the oracle's bit for bit:
- at every bucket boundary b / K of the row guides and one ulp either side, paired with rows
  across the table;
- at the marginal's boundaries (through the sampler's (x0 - a) / (1 - a) remap, with both float
  neighbours of each preimage), at 0 and just below 1, in the uniform-sphere branch, and at random
  points;
- over the baked procedural sky on both charts (plain and MIS-compensated, whose CDFs hold long
  flat runs) and registered pictures with black rows and columns (zero-mass rows and flat runs
  inside rows) and hot texels, at sizes whose K exceeds the CDF length (non-powers of two)."""
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, compiler, scenes
from pathtracerglsl_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(chart, compensated):
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural", "intensity": 1.0, "chart": chart, "compensated": compensated}
    return sc


def _picture(w, h, seed):
    """A picture whose CDFs have flat runs: two black bands of rows, black columns, hot texels."""
    rng = np.random.default_rng(seed)
    rgb = rng.uniform(0.1, 1.0, size=(h, w, 3)).astype(np.float32)
    rgb[: h // 6] = 0.0
    rgb[h // 2: h // 2 + 3] = 0.0
    rgb[:, w // 5: w // 5 + w // 4] = 0.0
    rgb[h // 3, (2 * w) // 3] = [60.0, 50.0, 40.0]
    rgb[(3 * h) // 4, w - 1] = [9.0, 9.0, 30.0]
    return rgb


def _k(n):
    k = 1
    while k < n:
        k <<= 1
    return k


def _inputs(w, h, a, seed):
    """(x0, x1) pairs: row-guide boundaries, marginal-guide boundaries, edges and random points."""
    rng = np.random.default_rng(seed)
    f = np.float32
    kc, km = _k(w), _k(h)
    b1 = (np.arange(kc + 1, dtype=np.float64) / kc).astype(f)
    x1 = np.concatenate([b1, np.nextafter(b1, f(-1)), np.nextafter(b1, f(2))])
    x1 = x1[(x1 >= 0) & (x1 < 1)]
    # marginal boundaries u = b / km reached through x0 = a + u (1 - a), with both neighbours
    u = np.arange(km + 1, dtype=np.float64) / km
    x0b = (a + u * (1.0 - a)).astype(f)
    x0 = np.concatenate([x0b, np.nextafter(x0b, f(-1)), np.nextafter(x0b, f(2))])
    x0 = x0[(x0 >= 0) & (x0 < 1)]
    pairs = [np.stack([np.repeat(rng.uniform(a, 1.0, 12).astype(f), len(x1)), np.tile(x1, 12)], 1),
             np.stack([np.repeat(x0, 4), rng.uniform(0, 1, 4 * len(x0)).astype(f)], 1),
             np.array([[0, 0], [0, np.nextafter(f(1), f(0))], [np.nextafter(f(1), f(0)), 0],
                       [np.nextafter(f(1), f(0))] * 2, [a, 0.5], [a * 0.5, 0.25]], f),
             rng.uniform(0, 1, size=(60000, 2)).astype(f)]
    return np.ascontiguousarray(np.concatenate(pairs).astype(f))


def _compare(engine, ors, xi):
    n = len(xi)
    got = np.empty((n, 4), np.float32)
    _lib.check(_lib.load().pt_debug_env_sample(engine.ctx, xi.ctypes.data, n, got.ctypes.data))
    dirs = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    ors.lib.or_env_sample(ors.h, xi.ctypes.data, n, dirs.ctypes.data, pdf.ctypes.data)
    diff = np.nonzero(~(got[:, :3] == dirs).all(1) | ~(got[:, 3] == pdf))[0]
    assert np.isfinite(got).all()
    assert len(diff) == 0, f"{len(diff)} of {n} differ, first at xi {xi[diff[:3]]}"


@pytest.mark.parametrize("chart,compensated", [("equirect", False), ("equirect", True), ("octahedral", False),
                                                ("octahedral", True)])
def test_guided_search_on_the_baked_sky_equals_the_oracle(oracle, chart, compensated):
    _say(f"[gpu] env guide, baked {chart} comp={compensated}")
    e = Engine(0)
    try:
        comp = compiler.compile_scene(_scene(chart, compensated), compiler.RenderStrategy("mis", 4, 3), 16, 16)
        e.load_renderer(comp)
        ors = oracle_lib.OracleScene(oracle, comp.scene)
        a = 0.0 if compensated else 0.0625
        h = comp.scene.env.table_h
        w = h if chart == "octahedral" else comp.scene.env.table_w     # the octahedral table is square
        _compare(e, ors, _inputs(w, h, a, 5))
    finally:
        e.close()


@pytest.mark.parametrize("chart,compensated,w,h", [("equirect", False, 96, 48), ("equirect", True, 300, 150),
                                                   ("octahedral", False, 40, 40), ("octahedral", True, 64, 64)])
def test_guided_search_on_registered_pictures_equals_the_oracle(oracle, chart, compensated, w, h):
    _say(f"[gpu] env guide, registered {chart} comp={compensated} {w}x{h}")
    e = Engine(0)
    try:
        comp = compiler.compile_scene(_scene(chart, False), compiler.RenderStrategy("mis", 4, 3), 16, 16)
        e.load_renderer(comp)
        rgb = _picture(w, h, 17)
        e.register_environment_table(rgb, w, h, chart=chart, compensation=compensated)
        ors = oracle_lib.OracleScene(oracle, comp.scene)
        ors.register_env_table(rgb, w, h, chart=compiler.ENV_CHARTS[chart], compensated=compensated)
        a = 0.0 if compensated else 0.0625
        _compare(e, ors, _inputs(w, h, a, 9))
    finally:
        e.close()
