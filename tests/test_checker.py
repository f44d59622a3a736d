"""The checker material (materials/checker/checker.glsl, checker.ts) and Hit.uv (glsl/core/structs.glsl:
40-43) on the oracle and the lowering.

Pins: under a white sky, pt with one bounce makes every pixel of a plane the albedo of the cell
its hit lies in (to an ulp); a quad whose visible hits all fall in cell (0, 0) renders albedo_a at every
pixel, one shifted a cell along u renders albedo_b (quad_uv's u axis, floor and the mod-2 parity),
both in RGB and spectral programs within the lift; with the parity inverted (a planted mistake in
the oracle) both fail; a checker sphere and mesh in a lit room render the mean of the two albedos'
pictures in between (the chart switches cells); the lowering accepts the checker maps that need no
other missing component, and materialsReadUv only with a checker."""
import json
import os

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, compiler, scenes

A, B = np.array([0.9, 0.6, 0.3], np.float32), np.array([0.1, 0.2, 0.4], np.float32)


def _film(lib, sc, color="rgb", spp=4, w=24, h=24):
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("none", 1, -1, color=color), w, h)
    f, _, _ = oracle_lib.OracleScene(lib, comp.scene).render(comp.program, comp.params(), spp, keep_samples=True)
    return f[..., :3]


@pytest.mark.parametrize("corner,want", [((-0.5, 0.0, 0.5), A), ((-1.5, 0.0, 0.5), B)])
def test_cells_render_their_albedo_exactly(oracle, corner, want):
    f = _film(oracle, scenes.checker_plane(corner=corner))
    assert np.allclose(f.reshape(-1, 3), np.broadcast_to(want, (f.shape[0] * f.shape[1], 3)), rtol=2e-7, atol=0)
    s = _film(oracle, scenes.checker_plane(corner=corner), "spectral", spp=256).reshape(-1, 3).mean(0)
    assert np.all(np.abs(s - want) < 0.03), s
    plants = oracle_lib.load("_plants")
    plants.or_set_plant(13)                    # the mod-2 parity inverted
    try:
        g = _film(plants, scenes.checker_plane(corner=corner))
        assert not np.allclose(g.reshape(-1, 3).mean(0), want, atol=0.05)
    finally:
        plants.or_set_plant(0)


def test_sphere_and_mesh_charts_switch_cells(oracle):
    comp = compiler.compile_scene(scenes.checker_room(), compiler.RenderStrategy("mis", 4, 3), 48, 48)
    f, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render(comp.program, comp.params(), 16, keep_samples=True)
    assert np.isfinite(f).all() and f[..., :3].mean() > 0.01


def test_lowering_accepts_checker_maps():
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "program_descriptions.json")))
    ok = set()
    for name, d in fix["cases"].items():
        if "checker" in d["materials"]["models"]:
            try:
                compiler.program_from_description(d, 16, 16, 1)
                ok.add(name)
            except _lib.PtError as e:
                assert "GRIN" in str(e), (name, str(e))
    assert {"charts + pt", "hearth + pt-nee", "spectral-materials + mis"} <= ok
    with pytest.raises(_lib.PtError, match="materialsReadUv"):
        compiler.program_from_description(fix["cases"]["expr-const + pt-nee"], 16, 16, 1)
