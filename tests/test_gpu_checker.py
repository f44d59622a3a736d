"""The checker material on the GPU: films equal to the oracle's bit for bit (the quad, sphere and
planar mesh charts, RGB and spectral, nee and mis, the albedo guide), and the frozen 'hearth +
pt-nee' link map lowered by the library."""
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import compiler, scenes

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "program_descriptions.json")))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("est,color", [("nee", "rgb"), ("mis", "rgb"), ("mis", "spectral"), ("none", "rgb")])
def test_checker_films_equal_the_oracle(engine, oracle, est, color):
    for name, sc in (("room", scenes.checker_room()), ("plane", scenes.checker_plane(corner=(-1.5, 0.0, 0.5)))):
        st = compiler.RenderStrategy(est, 5, 3, frames_in_flight=4, color=color)
        comp = compiler.compile_scene(sc, st, 32, 24)
        _say(f"[gpu] checker {name} {est} {color}")
        engine.load_renderer(comp)
        engine.clear_accumulation()
        engine.render_frames(8)
        g = engine.read_export()[..., :3]
        o, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render(comp.program, comp.params(), 8, keep_samples=True)
        assert np.array_equal(g, o[..., :3]), f"{name}: max diff {np.abs(g - o[..., :3]).max()}"


def test_checker_albedo_guide_equals_the_oracle(engine, oracle):
    st = compiler.RenderStrategy("nee", 4, 3, frames_in_flight=3, aovs=("albedo", "normal"))
    comp = compiler.compile_scene(scenes.checker_room(), st, 24, 24)
    _say("[gpu] checker albedo guide")
    engine.load_renderer(comp)
    engine.clear_accumulation()
    engine.render_frames(4)
    ga = engine.read_export("albedo")
    _, oa, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render_aovs(comp.program, comp.params(), 4)
    assert np.array_equal(ga[..., :3], oa[..., :3])


def test_frozen_checker_record_renders_like_the_oracle(engine, oracle):
    d = FIX["cases"]["hearth + pt-nee"]
    comp = compiler.compile_scene(scenes.checker_room(), compiler.RenderStrategy("nee", d["measurement"]["maxBounces"], 3,
                                                                                 frames_in_flight=4), 24, 16)
    _say("[gpu] description 'hearth + pt-nee' through pt_renderer_load_description")
    engine.load_renderer_description(d, comp)
    engine.clear_accumulation()
    engine.render_frames(6)
    g = engine.read_export()[..., :3]
    prog = compiler.program_from_description(d, comp.width, comp.height, comp.program.frames_in_flight)
    o, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render(prog, comp.params(), 6, keep_samples=True)
    assert np.array_equal(g, o[..., :3])
