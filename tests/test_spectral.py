"""Spectral color model: the host lift and Planck tables against the reference's exact witness
quantities, and the oracle's four-wavelength walk against the spectral witnesses.

Pins (all values from the reference's own test files):
- lift + CIE observer: spectral-plane's projected lifted albedo (0.19988754650944823, ...)
  (tests/witnesses/scenes/spectralFoundation.ts:48), a 0.25 nm quadrature of the float32 lift;
- coefficient lift + Beer-Lambert: spectral-slab / spectral-shadow-slab RGB
  (tests/witnesses/scenes/spectralMedia.ts:21-22) at distances 1 and 2;
- Planck normalisation: blackbody 1000 K / 40000 K channel ratios and unit reference Y
  (tests/components/blackbody.test.ts:10-50);
- the walk: spectral-direct = 1, spectral furnace 0.4 (1 - 0.5^7), coloured furnace and
  spectral slab within the witness tolerances (spectralFoundation.ts:33-48, spectralMedia.ts:41-43).
"""
import math

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import abi, compiler, scenes, spectral


def _coeffs(rgb, role):
    rec = spectral.spectrum_record(rgb, role)
    assert rec["kind"] == abi.SPEC_SIGMOID
    return [float(c) for c in rec["c"]], float(rec["scale"])


def test_curves_file_is_the_pinned_asset():
    c = spectral.curves_texture()
    assert c.shape == (2, 128, 4)
    z = spectral.z_nodes()
    assert z[0] == 0.0 and z[63] == 1.0 and np.all(np.diff(z) > 0)


def test_lifted_albedo_projects_to_the_spectral_plane_witness():
    c, _ = _coeffs([0.2, 0.5, 0.8], spectral.ROLE_REFLECTANCE)
    rgb = spectral.project_rgb(lambda l: spectral.sigmoid(c, l) * spectral.curves_at(l)[3])
    np.testing.assert_allclose(rgb, [0.19988754650944823, 0.49999884784204934, 0.8003487614363121], atol=2e-7)


@pytest.mark.parametrize("dist,expect", [
    (1.0, [0.7252018857862556, 0.49546203867937083, 0.2958918635233839]),
    (2.0, [0.5095338330733582, 0.24475284573080522, 0.08187756456891372]),
])
def test_coefficient_lift_beer_lambert(dist, expect):
    c, scale = _coeffs([0.3, 0.7, 1.2], spectral.ROLE_COEFFICIENT)
    assert scale == np.float32(2.4)
    rgb = spectral.project_rgb(lambda l: math.exp(-scale * spectral.sigmoid(c, l) * dist) * spectral.curves_at(l)[3])
    np.testing.assert_allclose(rgb, expect, atol=2e-7)


def test_grey_values_skip_the_sigmoid():
    r = spectral.spectrum_record(0.73, spectral.ROLE_REFLECTANCE)
    assert r["kind"] == abi.SPEC_GREY and r["value"] == np.float32(0.73)
    r = spectral.spectrum_record([12.0, 12.0, 12.0], spectral.ROLE_EMISSION)
    assert r["kind"] == abi.SPEC_GREY and r["scale"] == 24.0 and r["value"] == 0.5 and r["d65"] == 1
    assert spectral.spectrum_record(0.0, spectral.ROLE_COEFFICIENT)["kind"] == abi.SPEC_ZERO


def test_blackbody_projection_matches_reference_ratios():
    warm, cool = spectral.blackbody_color(1000.0)[0], spectral.blackbody_color(40000.0)[0]
    assert abs(warm[1] / warm[0] - 0.008696905553) < 1e-9
    assert abs(warm[2] / warm[0] - (-0.019583581807)) < 1e-9
    assert abs(cool[0] / cool[2] - 0.342393) < 1e-6
    for k in (1000, 3200, 6500, 12000, 40000):
        rgb, _ = spectral.blackbody_color(float(k))
        y = sum(spectral.RGB_TO_XYZ[1][i] * rgb[i] for i in range(3))
        assert abs(y - 1.0) < 1e-9
    assert min(spectral.blackbody_color(6500.0)[0]) > 0.85


def test_any_colour_lifts_and_a_tampered_atlas_is_refused(monkeypatch, tmp_path):
    sc = scenes.spectral_furnace((0.31, 0.17, 0.66))
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("none", 2, -1, color="spectral"), 8, 8)
    assert any(int(m["s_albedo"]["kind"]) == abi.SPEC_SIGMOID for m in comp.scene.materials)
    a = spectral.atlas().copy()
    a[12345, 1] = np.nextafter(a[12345, 1], np.float32(np.inf))
    bad = tmp_path / "atlas.zlib"
    spectral.write_atlas(a, str(bad))
    monkeypatch.setattr(spectral, "ATLAS_PATH", str(bad))
    monkeypatch.setattr(spectral, "_atlas", None)
    with pytest.raises(spectral.SpectralDataError, match="SHA-256"):
        spectral.atlas()


def test_spectral_procedural_sky_carries_the_atlas():
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural"}
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 2, -1, color="spectral"), 8, 8)
    assert comp.scene.spectral_atlas.shape == (3 * 64 ** 3, 4) and comp.scene.spectral_z_nodes.shape == (64,)
    d = comp.scene.desc()
    assert d.spectral_atlas and d.spectral_z_nodes
    rgb = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 2, -1), 8, 8)
    assert rgb.scene.spectral_atlas is None and not rgb.scene.desc().spectral_atlas


def _render(oracle, sc, strategy, w, h, spp):
    comp = compiler.compile_scene(sc, strategy, w, h)
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    film, _, _ = ors.render(comp.program, comp.params(), spp)
    return film[..., :3].reshape(-1, 3).mean(axis=0)


@pytest.mark.parametrize("wl", ["uniform", "visible"])
def test_spectral_direct_d65_white(oracle, wl):
    m = _render(oracle, scenes.spectral_direct(), compiler.RenderStrategy("none", 0, -1, color="spectral", wavelengths=wl),
                32, 32, 256)
    np.testing.assert_allclose(m, [1.0, 1.0, 1.0], atol=0.004)


@pytest.mark.parametrize("wl", ["uniform", "visible"])
def test_spectral_grey_furnace(oracle, wl):
    m = _render(oracle, scenes.spectral_furnace(), compiler.RenderStrategy("none", 6, -1, color="spectral", wavelengths=wl),
                32, 32, 256)
    np.testing.assert_allclose(m, 0.4 * (1 - 0.5 ** 7), atol=0.002)


@pytest.mark.parametrize("wl", ["uniform", "visible"])
def test_spectral_coloured_furnace(oracle, wl):
    m = _render(oracle, scenes.spectral_furnace((0.2, 0.5, 0.8)),
                compiler.RenderStrategy("none", 6, -1, color="spectral", wavelengths=wl), 32, 32, 256)
    np.testing.assert_allclose(m, [0.22947348594259948, 0.3932385892493329, 0.7652177863431524], atol=0.004)


@pytest.mark.parametrize("wl", ["uniform", "visible"])
def test_spectral_slab_beer_lambert(oracle, wl):
    # centre region only: the slab covers the central part of the view (camera at z = 2)
    comp = compiler.compile_scene(scenes.spectral_slab(), compiler.RenderStrategy("none", 1, -1, color="spectral",
                                                                                 wavelengths=wl), 24, 24)
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    film, _, _ = ors.render(comp.program, comp.params(), 512)
    m = film[8:16, 8:16, :3].reshape(-1, 3).mean(axis=0)
    np.testing.assert_allclose(m, [0.7252018857862556, 0.49546203867937083, 0.2958918635233839], atol=0.005)


def test_spectral_estimators_agree(oracle):
    """pt == nee == mis in the blackbody-lit glass-medium Cornell box (statistical)."""
    means = {}
    for est in ("none", "nee", "mis"):
        means[est] = _render(oracle, scenes.cornell_spectral_medium(),
                             compiler.RenderStrategy(est, 8, 3, color="spectral"), 24, 24, 128)
    for est in ("nee", "mis"):
        np.testing.assert_allclose(means[est], means["none"], rtol=0.05)
