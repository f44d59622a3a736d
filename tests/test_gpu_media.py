"""Participating media on the HIP path vs the CPU oracle, through the C ABI.

Scenes: a purely absorbing box behind a null interface (slab), an ambient scattering haze with
no surfaces (homogeneous medium everywhere), and a glass sphere filled with a scattering medium
inside the Cornell box (dielectric interface + medium NEE through the shadow walker). Same
pcg4d streams on both sides, so per-pixel radiance must agree within 1e-4 relative."""
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import compiler, scenes

pytestmark = pytest.mark.gpu
REL_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _render_both(engine, oracle, name, sc, strategy, w, h, spp):
    comp = compiler.compile_scene(sc, strategy, w, h)
    engine.load_renderer(comp)
    engine.clear_accumulation()
    _say(f"[gpu] render {name} {strategy.direct_lighting} {w}x{h}x{spp}")
    engine.render_frames(spp)
    g = engine.read_export()
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    o, _, _ = ors.render(comp.program, comp.params(), spp)
    d = np.abs(g[..., :3] - o[..., :3])
    rel = d / np.maximum(np.abs(o[..., :3]), 1e-6)
    exact = np.mean(np.all(g[..., :3] == o[..., :3], axis=-1))
    bad = np.mean(np.any(rel > REL_TOL, axis=-1))
    _say(f"[gpu] {name}: exact pixels {exact:.5f}, over 1e-4 rel {bad:.5f}, mean {o[..., :3].mean():.5f}")
    return g, o, bad


def test_slab_null_interface_matches_oracle(engine, oracle):
    _, o, bad = _render_both(engine, oracle, "slab", scenes.slab_witness(),
                             compiler.RenderStrategy("none", 4, -1), 24, 24, 16)
    assert o[..., :3].max() > 0.0
    assert bad == 0.0


@pytest.mark.parametrize("est", ["none", "mis"])
def test_ambient_haze_matches_oracle(engine, oracle, est):
    _, o, bad = _render_both(engine, oracle, "haze", scenes.furnace_scatter(),
                             compiler.RenderStrategy(est, 24, 3), 24, 24, 8)
    assert o[..., :3].max() > 0.0
    assert bad == 0.0


@pytest.mark.parametrize("est", ["mis", "nee", "none"])
def test_glass_medium_sphere_matches_oracle(engine, oracle, est):
    _, o, bad = _render_both(engine, oracle, "glass-medium", scenes.cornell_glass_medium(),
                             compiler.RenderStrategy(est, 16, 3), 32, 32, 8)
    assert o[..., :3].max() > 0.0
    assert bad == 0.0


def test_media_frames_in_flight_do_not_change_pixels(engine):
    sc = scenes.cornell_glass_medium()
    films = []
    for fif in (1, 3):
        comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 16, 3, 0.95, fif), 24, 20)
        engine.load_renderer(comp)
        engine.clear_accumulation()
        engine.render_frames(5)
        films.append(engine.read_export())
    assert np.array_equal(films[0], films[1])


@pytest.mark.parametrize("mls", ["vertex", "equiangular"])
def test_haze_medium_nee_placements_match_oracle(engine, oracle, mls):
    _, o, bad = _render_both(engine, oracle, f"haze-{mls}", scenes.haze(),
                             compiler.RenderStrategy("nee", 24, 4, medium_light_sampling=mls), 32, 24, 8)
    assert o[..., :3].max() > 0.0
    assert bad == 0.0


DT = dict(volume_sampling="delta-tracking")


@pytest.mark.parametrize("color", ["rgb", "spectral"])
def test_glowing_slab_absorb_arm_matches_oracle(engine, oracle, color):
    sc = scenes.emit_slab() if color == "rgb" else scenes.emit_slab(sigma_a=[0.3, 0.7, 1.2], emission=[0.4, 0.8, 1.2])
    _, o, bad = _render_both(engine, oracle, f"emit-slab-{color}", sc, compiler.RenderStrategy("none", 4, -1, color=color),
                             24, 24, 16)
    assert o[..., :3].max() > 1.0
    assert bad == 0.0


@pytest.mark.parametrize("bounces", [0, 2, 32])
def test_emit_sat_delta_tracking_matches_oracle(engine, oracle, bounces):
    _, o, bad = _render_both(engine, oracle, f"emit-sat-{bounces}", scenes.emit_sat(),
                             compiler.RenderStrategy("none", bounces, -1, **DT), 16, 16, 16)
    assert o[..., :3].mean() > 0.3
    assert bad == 0.0


@pytest.mark.parametrize("est", ["nee", "none"])
def test_emit_scatter_delta_tracking_with_nee_matches_oracle(engine, oracle, est):
    _, o, bad = _render_both(engine, oracle, f"emit-scatter-{est}", scenes.emit_scatter(),
                             compiler.RenderStrategy(est, 24, 4, **DT), 24, 18, 8)
    assert o[..., :3].mean() > 0.05
    assert bad == 0.0


def test_spectral_emit_scatter_matches_oracle(engine, oracle):
    sc = scenes.emit_scatter(sigma_a=(0.02, 0.05, 0.1), sigma_s=(0.2, 0.4, 0.7))
    _, o, bad = _render_both(engine, oracle, "spectral-emit-scatter", sc,
                             compiler.RenderStrategy("nee", 8, 3, color="spectral", **DT), 24, 18, 8)
    assert bad == 0.0


def test_delta_majorant_follows_material_updates(engine, oracle):
    """emit-driven (tests/witnesses/index.ts:2066-2078): raise sigma_a after upload; the library
    re-derives the majorant so the null coefficient stays >= 0 and the image matches the oracle
    compiled at the new values."""
    strat = compiler.RenderStrategy("none", 8, -1, **DT)
    comp = compiler.compile_scene(scenes.emit_sat(), strat, 16, 16)
    sc2 = scenes.emit_sat()
    sc2["materials"]["glow"]["medium"]["sigma_a"] = 3.5
    comp2 = compiler.compile_scene(sc2, strat, 16, 16)
    engine.load_renderer(comp)
    engine.update_materials(comp2.scene.materials)
    engine.clear_accumulation()
    engine.render_frames(8)
    g = engine.read_export()
    o, _, _ = oracle_lib.OracleScene(oracle, comp2.scene).render(comp2.program, comp2.params(), 8)
    rel = np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)
    assert (rel <= REL_TOL).all()


def test_orthographic_media_camera_inside_and_outside(engine, oracle):
    """Orthographic rays do not share an origin: each camera ray's medium is its own
    region_at (the glass-medium sphere straddles part of the film plane)."""
    sc = scenes.cornell_glass_medium()
    sc["camera"] = {"type": "orthographic", "scale": 0.9, "position": [0.0, 0.0, 0.3], "target": [0.0, 0.0, -2.0]}
    sc["id"] = "cornell-glass-medium-ortho-inside"
    _, o, bad = _render_both(engine, oracle, "glass-medium-ortho", sc, compiler.RenderStrategy("mis", 12, 3), 32, 32, 8)
    assert bad == 0.0


@pytest.mark.parametrize("est", ["mis", "none"])
def test_rough_glass_medium_matches_oracle(engine, oracle, est):
    """Rough glass bounding a scattering medium: the medium switch and eta_scale on the
    non-delta transmission lobe, NEE at the rough interface."""
    sc = scenes.cornell_rough_glass(0.25, medium={"sigma_s": [0.6, 0.8, 1.0], "sigma_a": [0.1, 0.1, 0.2], "phase_g": 0.3})
    _, o, bad = _render_both(engine, oracle, f"rough-glass-medium-{est}", sc, compiler.RenderStrategy(est, 12, 3), 32, 32, 8)
    assert o[..., :3].mean() > 0.01
    assert bad == 0.0


def test_thinlens_media_camera_straddles_the_medium(engine, oracle):
    """Thin-lens rays start over the lens disk (radius 0.3 at z = 0.45): the inner part of the
    disk lies inside the glass-medium sphere, the rim outside, so each camera ray takes its own
    region_at as its starting medium."""
    sc = scenes.cornell_glass_medium()
    sc["camera"] = {"type": "thinlens", "fov": 1.0, "aperture": 0.3, "focusDistance": 1.2,
                    "position": [0.0, 0.0, 0.45], "target": [0.0, 0.0, -2.0]}
    sc["id"] = "cornell-glass-medium-thinlens-inside"
    _, o, bad = _render_both(engine, oracle, "glass-medium-thinlens", sc, compiler.RenderStrategy("mis", 12, 3), 32, 32, 8)
    assert o[..., :3].mean() > 0.01
    assert bad == 0.0


@pytest.mark.parametrize("color", ["rgb", "spectral"])
def test_orthographic_slab_and_glow_match_oracle(engine, oracle, color):
    sc = scenes.orthographic(scenes.emit_slab() if color == "rgb" else
                             scenes.emit_slab(sigma_a=[0.3, 0.7, 1.2], emission=[0.4, 0.8, 1.2]))
    _, o, bad = _render_both(engine, oracle, f"ortho-glow-{color}", sc, compiler.RenderStrategy("none", 4, -1, color=color),
                             24, 24, 8)
    assert bad == 0.0


def _cornell_rayleigh_medium():
    sc = scenes.cornell_glass_medium()
    med = sc["materials"]["glass"]["medium"]
    sc["materials"]["glass"]["medium"] = {"model": "rayleigh", "sigma_s": med["sigma_s"], "sigma_a": med["sigma_a"]}
    return sc


@pytest.mark.parametrize("est", ["mis", "none"])
def test_rayleigh_glass_medium_matches_oracle(engine, oracle, est):
    # the Rayleigh phase (volume_scattering/rayleigh): medium NEE through its value and pdf, phase
    # sampling by its closed-form inversion
    _, o, bad = _render_both(engine, oracle, "rayleigh-glass-medium", _cornell_rayleigh_medium(),
                             compiler.RenderStrategy(est, 16, 3), 32, 32, 8)
    assert o[..., :3].max() > 0.0
    assert bad == 0.0


def test_rayleigh_furnace_matches_oracle(engine, oracle):
    _, o, bad = _render_both(engine, oracle, "rayleigh-furnace", scenes.rayleigh_furnace(),
                             compiler.RenderStrategy("none", 48, -1), 24, 24, 8)
    assert abs(o[..., :3].mean() - 0.4) < 0.03
    assert bad == 0.0
