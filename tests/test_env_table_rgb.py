"""The environment importance table rebuilt from a caller's RGB picture (registerEnvironmentTable,
src/engine/Engine.ts:450-460 -> buildEnvImportanceTable, src/components/env/importance.ts:16-127)
in the oracle (or_scene_env_table_rgb), pinned by the reference's own invariants
(tests/engine/buildEnvironmentSampler.test.ts): CDFs monotone and closing at 1 for a uniform
picture, an all-black picture giving total 0 without NaNs, the density read off CDF differences
summing to 1 and proportional to the blurred luminance times sin(theta), and the blur spreading a
lone bright texel to exactly its 8 neighbours (a far texel keeps density 0). Closed forms fix the
rest: a uniform picture's mass on the equirect chart is W sum_j sin(theta_j) times the
luminance weights' sum, the octahedral chart's CDFs of a uniform picture are (i + 1) / W, and MIS
compensation of a uniform picture leaves an all-zero table. The sampler then reads the table:
sampled directions' pdf equals environment_pdf and E[1/pdf] = 4 pi. Bad pictures (a negative or
non-finite sample, a non-square octahedral chart) are refused, as the reference throws."""
import math

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import abi, compiler, scenes

LUMA = 0.2126 + 0.7152 + 0.0722


def sky_oracle(oracle):
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural", "intensity": 1.0}
    comp = compiler.compile_scene(sc)
    return comp, oracle_lib.OracleScene(oracle, comp.scene)


def table(ors, w, h):
    return ors.env_table(w, h)


def density(cond, marg):
    dm = np.diff(np.concatenate([[0.0], marg.astype(np.float64)]))
    dc = np.diff(np.concatenate([np.zeros((cond.shape[0], 1)), cond.astype(np.float64)], axis=1), axis=1)
    return dc * dm[:, None]


def tent_blur(y, oct):
    """3 x 3 tent with weights (1 2 1) x (1 2 1) / 16: u wraps on the equirect chart and clamps
    on the octahedral one, v clamps."""
    h, w = y.shape
    out = np.zeros_like(y, dtype=np.float64)
    for dj, wj in ((-1, 1), (0, 2), (1, 1)):
        rows = np.clip(np.arange(h) + dj, 0, h - 1)
        for di, wi in ((-1, 1), (0, 2), (1, 1)):
            cols = np.clip(np.arange(w) + di, 0, w - 1) if oct else (np.arange(w) + di) % w
            out += wj * wi * y[rows][:, cols]
    return out / 16.0


def test_uniform_picture_closed_forms(oracle):
    comp, ors = sky_oracle(oracle)
    w, h = 16, 8
    tw = ors.register_env_table(np.full((h, w, 3), 1.0, np.float32), w, h)
    cond, marg = table(ors, w, h)
    sin_t = np.sin(np.pi * (np.arange(h) + 0.5) / h)
    assert tw == pytest.approx(w * LUMA * sin_t.sum(), rel=1e-12)
    assert np.all(np.diff(cond, axis=1) >= 0) and np.all(np.abs(cond[:, -1] - 1.0) <= 1e-6)
    assert np.all(np.diff(marg) >= 0) and abs(marg[-1] - 1.0) <= 1e-6
    np.testing.assert_allclose(cond, np.tile((np.arange(w) + 1) / w, (h, 1)), atol=1e-6)
    np.testing.assert_allclose(marg, np.cumsum(sin_t) / sin_t.sum(), atol=1e-6)
    # octahedral: equal-area texels, so a uniform picture's CDFs are linear in both directions
    n = 12
    tw = ors.register_env_table(np.full((n, n, 3), 0.5, np.float32), n, n, chart=abi.CHART_OCTAHEDRAL)
    cond, marg = table(ors, n, n)
    assert tw == pytest.approx(n * n * 0.5 * LUMA, rel=1e-12)
    np.testing.assert_allclose(cond, np.tile((np.arange(n) + 1) / n, (n, 1)), atol=1e-6)
    np.testing.assert_allclose(marg, (np.arange(n) + 1) / n, atol=1e-6)
    # MIS compensation removes the dOmega-weighted mean: a uniform picture leaves nothing
    tw = ors.register_env_table(np.full((h, w, 3), 2.0, np.float32), w, h, compensated=True)
    cond, marg = table(ors, w, h)
    assert tw == 0.0 and not cond.any() and not marg.any()


def test_black_picture_has_no_mass_and_no_nans(oracle):
    _, ors = sky_oracle(oracle)
    tw = ors.register_env_table(np.zeros((2, 2, 3), np.float32), 2, 2)
    cond, marg = table(ors, 2, 2)
    assert tw == 0.0
    assert np.isfinite(cond).all() and np.isfinite(marg).all() and not cond.any() and not marg.any()


@pytest.mark.parametrize("chart", [abi.CHART_EQUIRECT, abi.CHART_OCTAHEDRAL])
def test_density_from_cdf_differences(oracle, chart):
    _, ors = sky_oracle(oracle)
    w, h = (20, 9) if chart == abi.CHART_EQUIRECT else (10, 10)
    rgb = (((np.arange(w * h * 3) * 37) % 11) * 0.13).astype(np.float32).reshape(h, w, 3)
    tw = ors.register_env_table(rgb, w, h, chart=chart)
    cond, marg = table(ors, w, h)
    y = (0.2126 * rgb[..., 0].astype(np.float64) + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]).astype(np.float32)
    weight = tent_blur(y.astype(np.float64), chart == abi.CHART_OCTAHEDRAL)
    if chart == abi.CHART_EQUIRECT:
        weight = weight * np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    d = density(cond, marg)
    assert d.sum() == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_allclose(d, weight / weight.sum(), atol=2e-7)
    assert tw == pytest.approx(weight.sum(), rel=1e-6)


def test_blur_spreads_a_lone_texel_to_its_neighbours_only(oracle):
    _, ors = sky_oracle(oracle)
    w, h, ci, cj = 6, 5, 3, 2
    rgb = np.zeros((h, w, 3), np.float32)
    rgb[cj, ci, 0] = 100.0
    ors.register_env_table(rgb, w, h)
    d = density(*table(ors, w, h))
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            assert d[cj + dj, ci + di] > 0
    assert (d > 0).sum() == 9 and d[4, 0] == 0
    # at the phi seam the support wraps around (equirect u wraps)
    rgb[:] = 0
    rgb[cj, 0, 1] = 10.0
    ors.register_env_table(rgb, w, h)
    d = density(*table(ors, w, h))
    assert d[cj, w - 1] > 0 and d[cj, 1] > 0 and (d > 0).sum() == 9


def test_sampler_reads_the_registered_table(oracle):
    _, ors = sky_oracle(oracle)
    w, h = 64, 32
    rng = np.random.default_rng(5)
    rgb = rng.uniform(0.0, 1.0, size=(h, w, 3)).astype(np.float32)
    rgb[5, 40] = 400.0        # a hot spot the sampler must favour
    ors.register_env_table(rgb, w, h)
    n = 200000
    xi = rng.uniform(size=(n, 2)).astype(np.float32)
    dirs = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    oracle.or_env_sample(ors.h, xi.ctypes.data, n, dirs.ctypes.data, pdf.ctypes.data)
    pdf2 = np.zeros(n, np.float32)
    oracle.or_env_pdf(ors.h, dirs.ctypes.data, n, pdf2.ctypes.data)
    assert np.isclose(pdf, pdf2, rtol=1e-5).mean() > 0.995
    assert abs(np.mean(1.0 / pdf) / (4 * np.pi) - 1.0) < 0.02
    # the hot texel's cell (u = (40 + 0.5) / 64, v = (5 + 0.5) / 32) is drawn far above its area
    u = np.arctan2(dirs[:, 2], dirs[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(dirs[:, 1], -1, 1)) / np.pi
    hot = (np.floor(u * w) == 40) & (np.floor(v * h) == 5)
    area = (2 * np.pi / w) * (np.pi / h) * math.sin(np.pi * 5.5 / h) / (4 * np.pi)
    assert hot.mean() > 20 * area


def test_bad_pictures_are_refused(oracle):
    _, ors = sky_oracle(oracle)
    with pytest.raises(ValueError):
        ors.register_env_table(np.full((4, 4, 3), -1.0, np.float32), 4, 4)
    bad = np.ones((4, 4, 3), np.float32)
    bad[1, 1, 1] = np.nan
    with pytest.raises(ValueError):
        ors.register_env_table(bad, 4, 4)
    with pytest.raises(ValueError):
        ors.register_env_table(np.ones((4, 6, 3), np.float32), 6, 4, chart=abi.CHART_OCTAHEDRAL)
