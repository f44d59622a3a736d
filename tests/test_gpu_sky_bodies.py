"""Procedural sky bodies on the GPU (PT_SKY_*): films equal to the oracle's bit for bit for every
body in RGB and spectral programs on both charts (the proc-sky body's spectral lift through the
coefficient atlas, the grey bodies' value * D65), the frozen link maps of the four bodies lowered
by the library, the reference's witnesses rendered by the kernels (sky-missed-cap 0.095,
spectral-comp-empty 0.5, proc-sky-rotated nee == mis within 2 %), and a spectral program on the
coloured body refused when the scene carries no atlas."""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, abi, compiler, scenes

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "program_descriptions.json")))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gpu(engine, comp, spp):
    engine.load_renderer(comp)
    engine.clear_accumulation()
    engine.render_frames(spp)
    return engine.read_export()[..., :3]


def _oracle(oracle, comp, spp, prog=None):
    o, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render(prog or comp.program, comp.params(), spp,
                                                                keep_samples=True)
    return o[..., :3]


CASES = [("proc-sky", lambda ch: scenes.proc_sky(chart=ch), "mis", 4),
         ("proc-sky-rotated", scenes.proc_sky_rotated, "nee", 4),
         ("proc-sky-rotated", scenes.proc_sky_rotated, "mis", 4),
         ("sky-missed-cap", scenes.sky_missed_cap, "nee", 1),
         ("spectral-comp-empty", scenes.spectral_comp_empty, "mis", 1)]


@pytest.mark.parametrize("name,make,est,bounces", CASES, ids=[f"{c[0]}-{c[2]}" for c in CASES])
@pytest.mark.parametrize("color", ["rgb", "spectral"])
@pytest.mark.parametrize("chart", ["equirect", "octahedral"])
def test_sky_body_films_equal_the_oracle(engine, oracle, name, make, est, bounces, color, chart):
    st = compiler.RenderStrategy(est, bounces, 3 if bounces > 1 else -1, frames_in_flight=4, color=color)
    comp = compiler.compile_scene(make(chart), st, 24, 16)
    _say(f"[gpu] sky body {name} {est} {color} {chart}")
    g = _gpu(engine, comp, 8)
    o = _oracle(oracle, comp, 8)
    assert np.array_equal(g, o), f"max diff {np.abs(g - o).max()}"
    assert g.mean() > 0.01


def test_frozen_sky_body_records_render_like_the_oracle(engine, oracle):
    """Each of the four bodies' frozen link maps, lowered by pt_renderer_load_description."""
    cases = {"proc-sky-rotated + pt-mis": scenes.proc_sky_rotated("equirect"),
             "sky-missed-cap + spectral-octahedral": scenes.sky_missed_cap("octahedral"),
             "spectral-comp-empty + octahedral": scenes.spectral_comp_empty("octahedral"),
             "spectral-proc-sky + oct": scenes.proc_sky(chart="octahedral")}
    for name, sc in cases.items():
        d = FIX["cases"][name]
        color = d["measurement"]["color"]
        comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", d["measurement"]["maxBounces"], 3,
                                                                  frames_in_flight=4, color=color), 24, 16)
        _say(f"[gpu] description '{name}' through pt_renderer_load_description")
        engine.load_renderer_description(d, comp)
        engine.clear_accumulation()
        engine.render_frames(6)
        g = engine.read_export()[..., :3]
        prog = compiler.program_from_description(d, comp.width, comp.height, comp.program.frames_in_flight)
        assert np.array_equal(g, _oracle(oracle, comp, 6, prog)), name


@pytest.mark.parametrize("color", ["rgb", "spectral"])
@pytest.mark.parametrize("chart", ["equirect", "octahedral"])
def test_sky_missed_cap_witness(engine, color, chart):
    comp = compiler.compile_scene(scenes.sky_missed_cap(chart),
                                  compiler.RenderStrategy("nee", 1, -1, frames_in_flight=64, color=color), 32, 32)
    _say(f"[gpu] sky-missed-cap witness {color} {chart}")
    m = float(_gpu(engine, comp, 512).mean())
    assert abs(m - 0.095) < 0.004, m


@pytest.mark.parametrize("chart", ["equirect", "octahedral"])
def test_spectral_comp_empty_witness(engine, chart):
    comp = compiler.compile_scene(scenes.spectral_comp_empty(chart),
                                  compiler.RenderStrategy("mis", 1, -1, frames_in_flight=64, color="spectral"), 32, 32)
    _say(f"[gpu] spectral-comp-empty witness {chart}")
    m = float(_gpu(engine, comp, 512).mean())
    assert abs(m - 0.5) < 0.004, m


@pytest.mark.parametrize("color", ["rgb", "spectral"])
def test_proc_sky_rotated_witness(engine, color):
    means = []
    for est in ("nee", "mis"):
        comp = compiler.compile_scene(scenes.proc_sky_rotated(),
                                      compiler.RenderStrategy(est, 6, 3, frames_in_flight=64, color=color), 64, 48)
        _say(f"[gpu] proc-sky-rotated witness {est} {color}")
        means.append(float(_gpu(engine, comp, 192).mean()))
    assert abs(means[0] - means[1]) / means[0] < 0.02, means


def test_coloured_body_needs_the_atlas(engine):
    lib = _lib.load()
    comp = compiler.compile_scene(scenes.proc_sky(), compiler.RenderStrategy("mis", 2, -1, color="spectral"), 16, 8)
    comp.scene.spectral_atlas = None
    comp.scene.spectral_z_nodes = None
    engine.upload_scene(comp.scene)
    r = C.c_void_p()
    assert lib.pt_renderer_load(engine.ctx, C.byref(comp.program), C.byref(r)) == 2      # PT_E_MISSING_INPUT
    assert b"spectral_atlas" in lib.pt_last_error()
    grey = compiler.compile_scene(scenes.sky_missed_cap(), compiler.RenderStrategy("nee", 1, -1, color="spectral"), 16, 8)
    assert grey.scene.spectral_atlas is None
    engine.load_renderer(grey)                      # grey bodies lift without the atlas
