"""The C ABI's program-vs-scene rules and the ProgramDescription path on the GPU.

- Renderers loaded from the reference's own plan records (pt_renderer_load_description over the
  frozen complete link maps 'proc-sky + pt-mis' and 'cornell + orthographic', and a lowered
  record naming the octahedral chart) render films equal to the oracle's, bit for bit.
- The scene-dependent checks run at every render: a scene uploaded after the renderer loaded
  that the program cannot sample (a procedural sky under a spectral program, a sky table on
  the other chart) makes pt_render_frames / pt_forward fail instead of rendering, and
  Engine.select_renderer puts the selected renderer's own scene back.
- pt_scene_upload rejects an unknown sky chart; pt_batch_create rejects a program whose
  envSampler.chart differs from its scene's sky table."""
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


def _sky_cornell(chart="equirect", position=(0.0, 0.0, 0.6)):
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural", "intensity": 1.0, "chart": chart}
    sc["camera"]["position"] = list(position)
    sc["id"] = sc["id"] + f"-sky-{chart}"
    return sc


def _render_description(engine, oracle, description, comp, frames=8):
    engine.load_renderer_description(description, comp)
    engine.clear_accumulation()
    engine.render_frames(frames)
    g = engine.read_export()
    prog = compiler.program_from_description(description, comp.width, comp.height, comp.program.frames_in_flight)
    o, _, counts = oracle_lib.OracleScene(oracle, comp.scene).render(prog, comp.params(), frames)
    return g, o, counts, prog


@pytest.mark.parametrize("record", ["proc-sky + pt-mis", "cornell + orthographic"])
def test_frozen_description_renders_like_the_oracle(engine, oracle, record):
    d = FIX["full"][record]
    cam = d["measurement"]["camera"]
    sc = scenes.cornell_box()
    if record == "proc-sky + pt-mis":
        sc["environment"] = {"type": "procedural", "intensity": 1.0}
        sc["camera"] = {"position": cam["position"], "target": cam["target"], "fov": cam["fov"]["default"]}
        strategy = compiler.RenderStrategy("mis", 6, 3)
    else:
        sc["camera"] = {"type": "orthographic", "position": cam["position"], "target": cam["target"],
                        "scale": cam["scale"]}
        strategy = compiler.RenderStrategy("nee", 10, 3)
    comp = compiler.compile_scene(sc, strategy, 40, 32)
    _say(f"[gpu] description '{record}' through pt_renderer_load_description")
    g, o, counts, prog = _render_description(engine, oracle, d, comp)
    # the record selects the program this backend's own plan would have chosen
    assert bytes(prog) == bytes(comp.program)
    assert o[..., :3].mean() > 0.02
    assert np.array_equal(g[..., :3], o[..., :3])
    assert engine.ray_counts() == (int(counts[0]), int(counts[1]))


def test_octahedral_description_renders_like_the_oracle_and_batch_rejects_the_other_chart(engine, oracle):
    comp = compiler.compile_scene(_sky_cornell("octahedral"), compiler.RenderStrategy("mis", 8, 3), 40, 32)
    d = compiler.describe(_sky_cornell("octahedral"), compiler.RenderStrategy("mis", 8, 3))
    assert d["estimator"]["envSampler"]["chart"] == "octahedral"
    _say("[gpu] lowered octahedral-chart description over the octahedral-sky Cornell box")
    g, o, counts, prog = _render_description(engine, oracle, d, comp)
    assert prog.env_chart == abi.CHART_OCTAHEDRAL + 1
    assert o[..., :3].mean() > 0.05 and np.array_equal(g[..., :3], o[..., :3])
    # a batch whose program names the equirect chart over the octahedral table
    bad = abi.PtProgram.from_buffer_copy(bytes(comp.program))
    bad.env_chart = abi.CHART_EQUIRECT + 1
    lib = _lib.load()
    descs = (abi.PtSceneDesc * 1)(comp.scene.desc())
    progs = (abi.PtProgram * 1)(bad)
    prms = (abi.PtParams * 1)(comp.params())
    h = C.c_void_p()
    st = lib.pt_batch_create(engine.ctx, 1, descs, progs, prms, C.byref(h))
    assert st == 1 and b"envSampler.chart" in lib.pt_last_error()


def test_scene_replaced_after_load_is_checked_at_render(engine):
    lib = _lib.load()
    # a spectral program, loaded over a scene without a sky; then a procedural-sky scene compiled
    # for an RGB program arrives, without the coefficient atlas the spectral lift of its body needs
    spec = compiler.compile_scene(scenes.spectral_direct(), compiler.RenderStrategy("nee", 4, -1, color="spectral"), 16, 16)
    engine.load_renderer(spec)
    engine.render_frames(1)
    sky = compiler.compile_scene(_sky_cornell(), compiler.RenderStrategy("mis", 4, -1), 16, 16)
    engine.upload_scene(sky.scene)
    h, _ = engine._renderers[engine._selected]
    prm = spec.params()
    assert lib.pt_render_frames(h, C.byref(prm), 1) == 2              # PT_E_MISSING_INPUT
    assert b"spectral_atlas" in lib.pt_last_error()
    assert lib.pt_forward(h, C.byref(prm), 1) == 2
    # a program naming the octahedral chart, loaded over an octahedral table; then an equirect table
    octa = compiler.compile_scene(_sky_cornell("octahedral"), compiler.RenderStrategy("mis", 4, -1), 16, 16)
    octa.program.env_chart = abi.CHART_OCTAHEDRAL + 1
    engine.load_renderer(octa)
    engine.render_frames(1)
    engine.upload_scene(sky.scene)
    h, _ = engine._renderers[engine._selected]
    prm = octa.params()
    assert lib.pt_render_frames(h, C.byref(prm), 1) == 1
    assert b"chart" in lib.pt_last_error()
    # selecting a renderer makes its own scene resident again
    engine.select_renderer(octa.id)
    engine.clear_accumulation()
    engine.render_frames(2)
    assert np.isfinite(engine.read_export()).all()


def test_select_renderer_renders_its_own_scene(engine):
    a = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("mis", 6, 3), 24, 24)
    b = compiler.compile_scene(_sky_cornell(), compiler.RenderStrategy("mis", 6, 3), 24, 24)
    engine.load_renderer(a)
    engine.clear_accumulation()
    engine.render_frames(4)
    alone = engine.read_export()
    engine.load_renderer(b)
    engine.render_frames(1)
    engine.select_renderer(a.id)
    engine.clear_accumulation()
    engine.render_frames(4)
    assert np.array_equal(engine.read_export(), alone)


def test_scene_upload_rejects_unknown_chart(engine):
    sky = compiler.compile_scene(_sky_cornell(), compiler.RenderStrategy("mis", 4, -1), 16, 16)
    desc = sky.scene.desc()
    desc.env.chart = 7
    lib = _lib.load()
    assert lib.pt_scene_upload(engine.ctx, C.byref(desc)) == 1
    assert b"pt_env.chart" in lib.pt_last_error()
    engine._scene = None


def test_uniform_program_needs_the_uniform_table(engine):
    """A program that selects lights uniformly over a scene whose table was built for power
    selection is rejected at load; over the uniform table it renders."""
    lib = _lib.load()
    power = compiler.compile_scene(scenes.cornell_multi_light(), compiler.RenderStrategy("mis", 4, 3), 16, 16)
    uni = compiler.compile_scene(scenes.cornell_multi_light(), compiler.RenderStrategy("mis", 4, 3, light_selection="uniform"),
                                 16, 16)
    engine.upload_scene(power.scene)
    h = C.c_void_p()
    assert lib.pt_renderer_load(engine.ctx, C.byref(uni.program), C.byref(h)) == 1
    assert b"uniform" in lib.pt_last_error()
    engine.load_renderer(uni)
    engine.render_frames(1)
    assert np.isfinite(engine.read_export()).all()


@pytest.mark.parametrize("record", ["spectral-gaussian-plane + spectral-mis", "spectral-rayleigh-furnace + visible"])
def test_frozen_rayleigh_and_gaussian_records_render_like_the_oracle(engine, oracle, record):
    # link maps naming media.models ['rayleigh'] and spectrumKinds ['gaussian'], lowered through
    # pt_renderer_load_description over the witness scenes they were planned for
    d = FIX["cases"][record]
    if record.startswith("spectral-gaussian-plane"):
        comp = compiler.compile_scene(scenes.gaussian_plane(), compiler.RenderStrategy("mis", 1, -1, color="spectral"), 24, 24)
    else:
        comp = compiler.compile_scene(scenes.rayleigh_furnace(), compiler.RenderStrategy("none", 48, -1, color="spectral"), 24, 24)
    _say(f"[gpu] description '{record}' through pt_renderer_load_description")
    g, o, counts, prog = _render_description(engine, oracle, d, comp)
    assert prog.color == abi.COLOR_SPECTRAL and prog.max_bounces == d["measurement"]["maxBounces"]
    assert o[..., :3].mean() > 0.02
    assert np.array_equal(g[..., :3], o[..., :3])
    assert engine.ray_counts() == (int(counts[0]), int(counts[1]))
