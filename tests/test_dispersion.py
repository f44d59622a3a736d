"""Dispersive dielectrics: Cauchy / Sellmeier indices read at the hero wavelength, and the
once-per-path collapse of the spectral throughput to the hero lane (pt.ts:440,466-475,
components/color/spectral/spectral.glsl:12-24, features/intersection.ts:1609-1649).

Pins:
- N-BK7 through the Sellmeier formula equals the published Schott indices n_F, n_d, n_C;
- the RGB colour model reads the index at 587.5618 nm (components/color/spectra/*: toRGB) and
  keeps no formula; a formula constant in wavelength folds to its value;
- spectral-dispersion-slab (tests/witnesses/scenes/spectralDispersion.ts:87-95): the 0.25 nm
  quadrature target (0.91574, 0.91323, 0.90889) within 0.012 — a second collapse would
  multiply it by 4, a missing one leaves lanes that followed the hero's direction;
- the same slab in N-BK7 against a quadrature of (1 - F(n(lambda)))^2 D65 computed here;
- without dispersion the collapse never fires: a constant-index prism renders exactly as the
  same prism authored with a plain number.
"""

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import abi, compiler, scenes, spectral


def _render(oracle, sc, strategy, w, h, spp):
    comp = compiler.compile_scene(sc, strategy, w, h)
    film, _, counts = oracle_lib.OracleScene(oracle, comp.scene).render(comp.program, comp.params(), spp)
    return comp, film, counts


def test_bk7_sellmeier_matches_published_indices():
    B, C = scenes.N_BK7["sellmeier"]["B"], scenes.N_BK7["sellmeier"]["C"]
    for lam, n in ((486.1327, 1.52238), (587.5618, 1.51680), (656.2725, 1.51432)):
        assert abs(compiler.sellmeier_index(B, C, lam) - n) < 6e-6


def test_rgb_reads_the_reference_index_and_constants_fold():
    rgb = compiler.compile_scene(scenes.dispersion_slab(), compiler.RenderStrategy("none", 2, -1), 8, 8)
    m = rgb.scene.materials[0]
    assert m["ior_kind"] == abi.IOR_CONSTANT
    assert m["ior"] == np.float32(1.5 + 0.01 * (1000.0 / 587.5618) ** 2)
    sp = compiler.compile_scene(scenes.dispersion_slab(), compiler.RenderStrategy("none", 2, -1, color="spectral"), 8, 8)
    assert sp.scene.materials[0]["ior_kind"] == abi.IOR_CAUCHY
    flat = compiler.compile_scene(scenes.dispersion_slab({"cauchy": [1.5, 0.0]}),
                                  compiler.RenderStrategy("none", 2, -1, color="spectral"), 8, 8)
    assert flat.scene.materials[0]["ior_kind"] == abi.IOR_CONSTANT and flat.scene.materials[0]["ior"] == np.float32(1.5)
    with pytest.raises(compiler.CompileError, match="Cauchy"):
        compiler.compile_scene(scenes.dispersion_slab({"cauchy": [1.5, -0.1]}))
    with pytest.raises(compiler.CompileError, match="poles"):
        compiler.compile_scene(scenes.dispersion_slab({"sellmeier": {"B": [1, 0, 0], "C": [0.25, 0, 0]}}))


@pytest.mark.parametrize("wl", ["visible", "uniform"])
def test_dispersion_slab_witness(oracle, wl):
    # one wavelength per path after the collapse: ~1M samples for a standard error near 0.0025
    _, film, _ = _render(oracle, scenes.dispersion_slab(), compiler.RenderStrategy("none", 2, -1, color="spectral",
                                                                                    wavelengths=wl), 32, 32, 1024)
    m = film[..., :3].reshape(-1, 3).mean(axis=0)
    np.testing.assert_allclose(m, [0.9157439775502264, 0.9132305300063444, 0.908891337430429], atol=0.012)


def test_bk7_slab_against_quadrature(oracle):
    B, C = scenes.N_BK7["sellmeier"]["B"], scenes.N_BK7["sellmeier"]["C"]

    def two_transmissions(lam):
        n = compiler.sellmeier_index(B, C, lam)
        F = ((n - 1.0) / (n + 1.0)) ** 2
        return (1.0 - F) ** 2 * spectral.curves_at(lam)[3]

    expect = spectral.project_rgb(two_transmissions)
    _, film, _ = _render(oracle, scenes.dispersion_slab(scenes.N_BK7), compiler.RenderStrategy("none", 2, -1, color="spectral"),
                         32, 32, 1024)
    m = film[..., :3].reshape(-1, 3).mean(axis=0)
    np.testing.assert_allclose(m, expect, atol=0.012)


def test_constant_index_never_collapses(oracle):
    strat = compiler.RenderStrategy("none", 12, 3, color="spectral")
    _, a, ca = _render(oracle, scenes.dispersion_prism({"cauchy": [1.5, 0.0]}), strat, 32, 24, 8)
    _, b, cb = _render(oracle, scenes.dispersion_prism(1.5), strat, 32, 24, 8)
    assert a[..., :3].mean() > 0.0
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


def test_dispersive_prism_differs_from_its_reference_index(oracle):
    """The dispersive prism converges to a different image than the constant n_d prism (the
    stripe splits); both carry the same total energy within noise."""
    strat = compiler.RenderStrategy("none", 12, 3, color="spectral")
    _, a, _ = _render(oracle, scenes.dispersion_prism(), strat, 32, 24, 64)
    _, b, _ = _render(oracle, scenes.dispersion_prism(compiler.cauchy_index([1.45, 0.02], compiler.INDEX_REFERENCE_NM)),
                      strat, 32, 24, 64)
    assert not np.array_equal(a, b)
    ma, mb = a[..., :3].mean(), b[..., :3].mean()
    assert abs(ma - mb) < 0.1 * mb, (ma, mb)
