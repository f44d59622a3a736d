"""pt_render_frames_noalloc (the allocation-free render entry point) against pt_render_frames,
through the C ABI: two calls in a row on different renderers, the second with a smaller film,
and a film at the array's full capacity that needs more waves than the allocating path; every
film bit-identical to the allocating path's. A film above the capacity is rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

from pathtracerglsl_amd import _lib, compiler, scenes

pytestmark = pytest.mark.gpu
MAX_PATHS = 262144


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(engine, comp, frames):
    """(allocating film, allocation-free film) of the same renderer and samples."""
    lib = _lib.load()
    engine.load_renderer(comp)
    h, _ = engine._renderers[engine._selected]
    prm = comp.params()
    engine.clear_accumulation()
    _lib.check(lib.pt_render_frames(h, C.byref(prm), frames))
    ref = engine.read_export()
    engine.clear_accumulation()
    _lib.check(lib.pt_render_frames_noalloc(h, C.byref(prm), frames))
    got = engine.read_export()
    assert engine.sample_count == frames
    return ref, got


def test_noalloc_equals_render_frames(engine):
    cases = [
        ("cloud 96x64 mis fif 4", compiler.compile_scene(scenes.icosphere_cloud(n_spheres=60, seed=3, box=25.0),
                                                          compiler.RenderStrategy("mis", 8, 3, 0.95, 4), 96, 64), 8),
        ("cornell multi-light 40x30 nee", compiler.compile_scene(scenes.cornell_multi_light(),
                                                                  compiler.RenderStrategy("nee", 6, 3), 40, 30), 5),
    ]
    for name, comp, frames in cases:
        _say(f"[gpu] noalloc vs render_frames: {name}")
        ref, got = _both(engine, comp, frames)
        assert ref[..., :3].mean() > 0.01
        assert np.array_equal(ref, got), name


def test_noalloc_at_capacity_and_over(engine):
    comp = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("mis", 8, 3, 0.95, 4), 512, 512)
    _say("[gpu] noalloc at capacity: 512x512 (one frame per wave) vs render_frames (four)")
    ref, got = _both(engine, comp, 4)
    assert np.array_equal(ref, got)
    big = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("mis", 4, 3), 1024, 512)
    engine.load_renderer(big)
    h, _ = engine._renderers[engine._selected]
    prm = big.params()
    lib = _lib.load()
    assert lib.pt_render_frames_noalloc(h, C.byref(prm), 1) == 1
    assert b"262144" in lib.pt_last_error()
