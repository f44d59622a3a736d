"""The batched persistent wavefront (pt_batch_*) against plain per-scene launches and the
oracle: every film of a mixed-scene, hipGraph-captured batch is bit-identical to pt_render_frames
of that scene alone, for several wave sizes (scenes split across waves), block sizes and stack
depths, with and without graph capture, and equal to the oracle's film of that scene; the
64-scene workload (log-uniform 1k-500k triangles, 128^2-1024^2, 4-256 spp) renders as one batch
identical to its 64 sequential launches, its cheapest scenes equal to the oracle's films; the
autotuner returns a valid configuration."""
import sys

import numpy as np
import pytest

import oracle_lib
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


def test_batch_films_equal_the_oracle(engine, oracle, small_batch):
    comps, spp = small_batch
    br = BatchRenderer(engine, comps)
    br.render(spp)
    for s, (c, n) in enumerate(zip(comps, spp)):
        g = br.read_export(s)
        o, _, _ = oracle_lib.OracleScene(oracle, c.scene).render(c.program, c.params(), n)
        same = np.array_equal(g[..., :3], o[..., :3])
        _say(f"[gpu] batch scene {s} ({c.width}x{c.height}x{n}) vs oracle: bit-identical {same}")
        assert same, s
    br.close()


def test_batch_64_scene_workload(engine, oracle):
    """The 64-scene batch workload at its own sizes and spp, pt-mis depth 12: one batch equals the
    64 sequential launches film for film; its four cheapest scenes equal the oracle's films."""
    items = scenes.batch_workload()
    st = compiler.RenderStrategy("mis", 12, 3, 0.95, 4)
    comps = [compiler.compile_scene(sc, st, w, h) for sc, w, h, _ in items]
    spp = [n for _, _, _, n in items]
    _say(f"[gpu] 64-scene batch: {sum(len(c.scene.tri_object) for c in comps)} triangles, "
         f"{sum(c.width * c.height * n for c, n in zip(comps, spp)) / 1e6:.1f} Mpaths")
    ref = _sequential(engine, comps, spp)
    _say("[gpu] 64-scene batch: sequential launches done")
    br = BatchRenderer(engine, comps)
    br.render(spp)
    diff = [s for s in range(len(comps)) if not np.array_equal(br.read_export(s), ref[s])]
    _say(f"[gpu] 64-scene batch: {len(comps) - len(diff)} of {len(comps)} films bit-identical to sequential")
    assert not diff
    assert all(np.isfinite(f).all() for f in ref)
    cost = sorted(range(len(comps)), key=lambda s: comps[s].width * comps[s].height * spp[s] * (1 + len(comps[s].scene.tri_object)) ** 0.25)
    for s in cost[:4]:
        c = comps[s]
        o, _, _ = oracle_lib.OracleScene(oracle, c.scene).render(c.program, c.params(), spp[s])
        same = np.array_equal(ref[s][..., :3], o[..., :3])
        _say(f"[gpu] 64-scene batch scene {s} ({c.width}x{c.height}x{spp[s]}, {len(c.scene.tri_object)} tris) vs oracle: {same}")
        assert same, s
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
