"""The oneshot accumulator on the GPU: the film is the latest sample, across waves and calls, equal
to the oracle's bit for bit; the guides follow it; the sample count advances; a oneshot program
with variance and a batch of oneshot programs are refused; the frozen 'cornell-oneshot' link map
renders through pt_renderer_load_description like the oracle."""
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


@pytest.mark.parametrize("est,color,fif", [("nee", "rgb", 2), ("mis", "spectral", 4)])
def test_oneshot_film_equals_the_oracle(engine, oracle, est, color, fif):
    st = compiler.RenderStrategy(est, 8, 3, frames_in_flight=fif, color=color, accumulation="oneshot")
    comp = compiler.compile_scene(scenes.cornell_box(), st, 32, 24)
    _say(f"[gpu] oneshot {est} {color} frames in flight {fif}")
    engine.load_renderer(comp)
    engine.clear_accumulation()
    engine.render_frames(3)
    engine.render_frames(2)
    g = engine.read_export()
    assert engine.sample_count == 5
    o, samples, _ = oracle_lib.OracleScene(oracle, comp.scene).render(comp.program, comp.params(), 5, keep_samples=True)
    assert np.array_equal(o[..., :3], samples[-1])
    assert np.array_equal(g[..., :3], o[..., :3])


def test_oneshot_guides_equal_the_oracle(engine, oracle):
    st = compiler.RenderStrategy("nee", 6, 3, frames_in_flight=3, aovs=("albedo", "normal"), accumulation="oneshot")
    comp = compiler.compile_scene(scenes.cornell_box(), st, 24, 24)
    engine.load_renderer(comp)
    engine.clear_accumulation()
    engine.render_frames(4)
    g, ga, gn = engine.read_export(), engine.read_export("albedo"), engine.read_export("normal")
    o, oa, on, _ = oracle_lib.OracleScene(oracle, comp.scene).render_aovs(comp.program, comp.params(), 4)
    assert np.array_equal(g[..., :3], o[..., :3])
    assert np.array_equal(ga[..., :3], oa[..., :3]) and np.array_equal(gn[..., :3], on[..., :3])


def test_oneshot_refusals(engine):
    lib = _lib.load()
    comp = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("nee", 4, 3, accumulation="oneshot"), 16, 16)
    bad = abi.PtProgram.from_buffer_copy(bytes(comp.program))
    bad.variance = 1
    r = C.c_void_p()
    assert lib.pt_renderer_load(engine.ctx, C.byref(bad), C.byref(r)) == 1
    assert b"oneshot" in lib.pt_last_error()
    descs = (abi.PtSceneDesc * 1)(comp.scene.desc())
    progs = (abi.PtProgram * 1)(comp.program)
    prms = (abi.PtParams * 1)(comp.params())
    h = C.c_void_p()
    assert lib.pt_batch_create(engine.ctx, 1, descs, progs, prms, C.byref(h)) == 5
    assert b"accumulation" in lib.pt_last_error()


def test_frozen_oneshot_record_renders_like_the_oracle(engine, oracle):
    d = FIX["cases"]["cornell-oneshot + oneshot"]
    comp = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("nee", 8, 3), 32, 24)
    _say("[gpu] description 'cornell-oneshot + oneshot' through pt_renderer_load_description")
    engine.load_renderer_description(d, comp)
    engine.clear_accumulation()
    engine.render_frames(3)
    g = engine.read_export()
    prog = compiler.program_from_description(d, comp.width, comp.height, comp.program.frames_in_flight)
    assert prog.accumulation == abi.ACC_ONESHOT
    o, samples, _ = oracle_lib.OracleScene(oracle, comp.scene).render(prog, comp.params(), 3, keep_samples=True)
    assert np.array_equal(o[..., :3], samples[-1]) and o[..., :3].mean() > 0.02
    assert np.array_equal(g[..., :3], o[..., :3])
