"""Light-tree selection on the GPU (PT_SEL_BVH): films equal to the oracle's bit for bit (RGB and
spectral, nee and mis, two-sided rough-glass queries, a medium's normal-free queries, an
environment taking half of the draws), the frozen 'two-light + bvh' link map lowered by the
library, the reference's two-light witness and the many-lights power == bvh checks rendered by
the kernels, and the refusals (no tree inputs, an environment probability other than 1/2,
equiangular medium sampling, the adjoint, batches)."""
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


def _sky(sc):
    sc = dict(sc)
    sc["environment"] = {"type": "procedural", "intensity": 0.5}
    return sc


CASES = [("many", lambda: scenes.many_lights(), "nee", "rgb"),
         ("many", lambda: scenes.many_lights(), "mis", "rgb"),
         ("many-glass", lambda: scenes.many_lights(rough_glass=True), "mis", "rgb"),
         ("many", lambda: scenes.many_lights(), "mis", "spectral"),
         ("two-light", scenes.two_light, "nee", "rgb"),
         ("many-sky", lambda: _sky(scenes.many_lights(16)), "mis", "rgb"),
         ("glass-medium", scenes.cornell_glass_medium, "mis", "rgb"),
         ("cornell-multi", scenes.cornell_multi_light, "mis", "spectral")]


@pytest.mark.parametrize("name,make,est,color", CASES, ids=[f"{c[0]}-{c[2]}-{c[3]}" for c in CASES])
def test_tree_films_equal_the_oracle(engine, oracle, name, make, est, color):
    st = compiler.RenderStrategy(est, 5, 3, frames_in_flight=4, color=color, light_selection="bvh")
    comp = compiler.compile_scene(make(), st, 32, 24)
    _say(f"[gpu] light tree {name} {est} {color}")
    g = _gpu(engine, comp, 8)
    o = _oracle(oracle, comp, 8)
    assert np.array_equal(g, o), f"max diff {np.abs(g - o).max()}"
    assert g.mean() > 1e-3


def test_frozen_two_light_bvh_record_renders_like_the_oracle(engine, oracle):
    d = FIX["cases"]["two-light + bvh"]
    comp = compiler.compile_scene(scenes.two_light(), compiler.RenderStrategy("nee", d["measurement"]["maxBounces"], 3,
                                                                              frames_in_flight=4), 32, 24)
    _say("[gpu] description 'two-light + bvh' through pt_renderer_load_description")
    engine.load_renderer_description(d, comp)
    engine.clear_accumulation()
    engine.render_frames(6)
    g = engine.read_export()[..., :3]
    prog = compiler.program_from_description(d, comp.width, comp.height, comp.program.frames_in_flight)
    assert prog.light_selection == abi.SEL_BVH
    assert np.array_equal(g, _oracle(oracle, comp, 6, prog))


@pytest.mark.parametrize("make,est", [(scenes.two_light, "nee"), (lambda: scenes.many_lights(), "nee"),
                                      (lambda: scenes.many_lights(), "mis")])
def test_power_equals_bvh_on_the_kernels(engine, make, est):
    means = {}
    for sel in ("power", "bvh"):
        comp = compiler.compile_scene(make(), compiler.RenderStrategy(est, 4, 3, frames_in_flight=64, light_selection=sel),
                                      64, 48)
        _say(f"[gpu] power == bvh {comp.id} {sel}")
        means[sel] = float(_gpu(engine, comp, 192).mean())
    assert abs(means["power"] - means["bvh"]) / means["power"] < 0.02, means


def test_refusals(engine):
    lib = _lib.load()
    comp = compiler.compile_scene(scenes.many_lights(8), compiler.RenderStrategy("mis", 4, 3, light_selection="bvh"), 16, 16)
    comp.scene.light_boxes = None
    comp.scene.light_powers = None
    engine.upload_scene(comp.scene)
    r = C.c_void_p()
    assert lib.pt_renderer_load(engine.ctx, C.byref(comp.program), C.byref(r)) == 2          # PT_E_MISSING_INPUT
    assert b"light_boxes" in lib.pt_last_error()
    sky = compiler.compile_scene(_sky(scenes.many_lights(8)), compiler.RenderStrategy("mis", 4, 3), 16, 16)
    engine.upload_scene(sky.scene)
    bvh = abi.PtProgram.from_buffer_copy(bytes(sky.program))
    bvh.light_selection = abi.SEL_BVH
    assert lib.pt_renderer_load(engine.ctx, C.byref(bvh), C.byref(r)) == 1                    # select_prob != 1/2
    assert b"1/2" in lib.pt_last_error()
    good = compiler.compile_scene(scenes.many_lights(8), compiler.RenderStrategy("nee", 4, 3, light_selection="bvh"), 16, 16)
    engine.upload_scene(good.scene)
    eq = abi.PtProgram.from_buffer_copy(bytes(good.program))
    eq.medium_light_sampling = abi.MLS_EQUIANGULAR
    assert lib.pt_renderer_load(engine.ctx, C.byref(eq), C.byref(r)) == 5                     # PT_E_UNSUPPORTED
    descs = (abi.PtSceneDesc * 1)(good.scene.desc())
    progs = (abi.PtProgram * 1)(good.program)
    prms = (abi.PtParams * 1)(good.params())
    h = C.c_void_p()
    assert lib.pt_batch_create(engine.ctx, 1, descs, progs, prms, C.byref(h)) == 5
    engine.load_renderer(good)
    hr, _ = engine._renderers[engine._selected]
    prm = good.params()
    ap = abi.PtAdjointParams()
    dimg = np.zeros((16, 16, 4), np.float32)
    dpar = np.zeros(64, np.float32)
    assert lib.pt_backward(hr, C.byref(prm), 1, C.byref(ap), dimg.ctypes.data, dpar.ctypes.data) == 5
    assert b"light tree" in lib.pt_last_error()
