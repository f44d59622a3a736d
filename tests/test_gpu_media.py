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
