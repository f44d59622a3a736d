"""The batched persistent wavefront (pt_batch_*) against plain per-scene launches: every film
of a mixed-scene, hipGraph-captured batch is bit-identical to pt_render_frames of that scene
alone, for several wave sizes (scenes split across waves), block sizes and stack depths, with
and without graph capture; the autotuner returns a valid configuration."""
import sys

import numpy as np
import pytest

from pathtracerglsl_amd import abi, compiler, scenes
from pathtracerglsl_amd.batch import BatchRenderer

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_batch():
    st = compiler.RenderStrategy("mis", 8, 3, 0.95, 2)
    items = [(scenes.cornell_box(), 40, 30, 5), (scenes.icosphere_cloud(n_spheres=40, seed=5, box=20.0), 64, 48, 3),
             (scenes.cornell_box((5.0, 9.0, 14.0)), 24, 24, 9), (scenes.icosphere_cloud(n_spheres=12, seed=6, box=8.0), 33, 17, 4),
             (scenes.icosphere_cloud(n_spheres=150, seed=8, box=30.0), 80, 60, 2)]
    comps = [compiler.compile_scene(sc, st, w, h) for sc, w, h, _ in items]
    return comps, [spp for _, _, _, spp in items]


def _sequential(engine, comps, spp):
    out = []
    for c, n in zip(comps, spp):
        engine.load_renderer(c)
        engine.clear_accumulation()
        engine.render_frames(n)
        out.append(engine.read_export())
    return out


def test_batch_equals_sequential_launches(engine, small_batch):
    comps, spp = small_batch
    ref = _sequential(engine, comps, spp)
    br = BatchRenderer(engine, comps)
    base = br.config
    _say(f"[gpu] batch default config {base.as_dict()}")
    variants = [dict(), dict(use_graph=0), dict(wave_paths=5000), dict(wave_paths=9600, trace_block=64, stack=32),
                dict(trace_block=256, trace_blocks_per_cu=4, shade_blocks_per_cu=4)]
    for v in variants:
        c = abi.PtBatchConfig()
        for k, val in base.as_dict().items():
            setattr(c, k, val)
        for k, val in v.items():
            setattr(c, k, val)
        br.config = c
        br.render(spp)
        for s in range(len(comps)):
            g = br.read_export(s)
            assert np.array_equal(g, ref[s]), (v, s, float(np.abs(g - ref[s]).max()))
        _say(f"[gpu] batch {v or 'default'}: all {len(comps)} films bit-identical")
    br.render(spp)          # replay of the cached graph
    assert all(np.array_equal(br.read_export(s), ref[s]) for s in range(len(comps)))
    br.close()


def test_batch_autotune(engine, small_batch):
    comps, spp = small_batch
    br = BatchRenderer(engine, comps)
    best, sweep = br.autotune(spp, probe_div=1)
    _say(f"[gpu] autotune best {best.as_dict()} over {len(sweep)} configs")
    assert len(sweep) >= 6 and best.trace_block in (64, 128, 256)
    br.render(spp)
    ref = _sequential(engine, comps, spp)
    assert all(np.array_equal(br.read_export(s), ref[s]) for s in range(len(comps)))
    br.close()
