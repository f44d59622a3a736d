"""Image environments on the GPU (PT_ENV_IMAGE): films equal to the oracle's bit for bit (RGB and
spectral through the coefficient atlas, both charts, compensated tables, rotations, with quad
lights sharing the draws), the frozen spectral-image-constant link maps lowered by the library,
and the witness value (the plane renders the picture's RGB) on the kernels."""
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


def _gpu(engine, comp, spp):
    engine.load_renderer(comp)
    engine.clear_accumulation()
    engine.render_frames(spp)
    return engine.read_export()[..., :3]


def _oracle(oracle, comp, spp, prog=None):
    o, _, _ = oracle_lib.OracleScene(oracle, comp.scene).render(prog or comp.program, comp.params(), spp,
                                                                keep_samples=True)
    return o[..., :3]


def _cornell_image(chart, comp, rot):
    sc = scenes.cornell_box()
    sc["environment"] = scenes.image_plane(scenes.sun_picture(64, 32), intensity=1.2, rotation=rot, chart=chart,
                                           compensated=comp, table_n=40)["environment"]
    return sc


CASES = [("equirect", False, 0.0, "mis", "rgb"), ("equirect", True, -2.0, "mis", "rgb"),
         ("octahedral", False, 0.73, "nee", "rgb"), ("octahedral", True, 1.1, "mis", "spectral"),
         ("equirect", False, 2.5, "mis", "spectral")]


@pytest.mark.parametrize("chart,comp,rot,est,color", CASES)
def test_image_films_equal_the_oracle(engine, oracle, chart, comp, rot, est, color):
    for name, sc in (("plane", scenes.image_plane(scenes.sun_picture(64, 32), rotation=rot, chart=chart, compensated=comp,
                                                  table_n=40)),
                     ("cornell", _cornell_image(chart, comp, rot))):
        st = compiler.RenderStrategy(est, 4, 3, frames_in_flight=4, color=color)
        c = compiler.compile_scene(sc, st, 24, 16)
        _say(f"[gpu] image env {name} {chart} comp={comp} rot={rot} {est} {color}")
        g = _gpu(engine, c, 8)
        o = _oracle(oracle, c, 8)
        assert np.array_equal(g, o), f"{name}: max diff {np.abs(g - o).max()}"
        assert g.mean() > 1e-3


def test_frozen_image_records_render_like_the_oracle(engine, oracle):
    names = [n for n in FIX["cases"] if n.startswith("spectral-image-constant +")]
    assert names
    for name in names:
        d = FIX["cases"][name]
        chart = (d["estimator"].get("envSampler") or {}).get("chart", "equirect")
        comp = compiler.compile_scene(scenes.image_plane(chart=chart),
                                      compiler.RenderStrategy("mis", d["measurement"]["maxBounces"], 3, frames_in_flight=4,
                                                              color=d["measurement"]["color"]), 24, 16)
        _say(f"[gpu] description '{name}' through pt_renderer_load_description")
        engine.load_renderer_description(d, comp)
        engine.clear_accumulation()
        engine.render_frames(6)
        g = engine.read_export()[..., :3]
        prog = compiler.program_from_description(d, comp.width, comp.height, comp.program.frames_in_flight)
        assert np.array_equal(g, _oracle(oracle, comp, 6, prog)), name


def test_image_constant_witness_on_the_kernels(engine):
    comp = compiler.compile_scene(scenes.image_plane(), compiler.RenderStrategy("nee", 1, -1, frames_in_flight=64), 32, 32)
    _say("[gpu] spectral-image-constant witness (RGB nee)")
    m = _gpu(engine, comp, 512).reshape(-1, 3).mean(0)
    assert np.all(np.abs(m - np.array(scenes.IMAGE_CONSTANT_RGB)) < 0.008), m
