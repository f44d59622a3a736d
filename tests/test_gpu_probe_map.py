"""The containment probe on the GPU through the probe map (scene_region_at,
features/intersection.ts:1121-1265, with mesh_inside_bvh's fixed-direction ray,
intersection/mesh/mesh.glsl:202-296): the innermost region at a point equals the oracle's, whose
probe walks the object's BVH, on points built to stress the map's conservative footprints:

- the classify offsets p +- 1e-3 n of random surface points (the probes every dielectric hit
  runs), with triangles nearly parallel to the probe direction oversampled (their projections are
  slivers, where the Moller-Trumbore test's rounding is largest);
- surface points themselves, mesh vertices and edge midpoints (rays through shared edges and
  vertices, where several triangles tie or crack);
- points a few ulps off the surface, and random points across the scene;
- the interpenetrating dense cloud (points inside other closed meshes) and an instance batch
  (the probe in each prototype's frame).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, compiler, scenes
from pathtracerglsl_amd.engine import Engine

pytestmark = pytest.mark.gpu
PROBE_DIR = np.array([0.5320544, 0.7396997, 0.4114743], np.float32)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _surface_points(tris, n, rng):
    """Probe points on and around random triangles of tris [m][9] (f32)."""
    a, b, c = tris[:, 0:3].astype(np.float64), tris[:, 3:6].astype(np.float64), tris[:, 6:9].astype(np.float64)
    g = np.cross(b - a, c - a)
    area = np.linalg.norm(g, axis=1)
    nh = g / np.maximum(area, 1e-30)[:, None]
    # half uniform over triangles, half over the triangles nearly parallel to the probe direction
    sliver = np.nonzero(np.abs(nh @ (PROBE_DIR / np.linalg.norm(PROBE_DIR))) < 0.05)[0]
    idx = np.concatenate([rng.integers(0, len(tris), n // 2),
                          sliver[rng.integers(0, len(sliver), n - n // 2)] if len(sliver) else rng.integers(0, len(tris), n - n // 2)])
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    bary = np.stack([1 - s, s * (1 - r2), s * r2], axis=1)
    kind = rng.integers(0, 6, n)
    bary[kind == 3] = np.eye(3)[rng.integers(0, 3, (kind == 3).sum())]                     # vertices
    e = rng.integers(0, 3, (kind == 4).sum())
    mid = np.full(((kind == 4).sum(), 3), 0.5)
    mid[np.arange(len(e)), e] = 0.0
    bary[kind == 4] = mid                                                                   # edge midpoints
    p = (bary[:, 0:1] * a[idx] + bary[:, 1:2] * b[idx] + bary[:, 2:3] * c[idx]).astype(np.float32)
    nn = nh[idx].astype(np.float32)
    out = p.copy()
    out[kind == 0] = p[kind == 0] + nn[kind == 0] * np.float32(1e-3)        # the classify offsets
    out[kind == 1] = p[kind == 1] - nn[kind == 1] * np.float32(1e-3)
    ulp = np.spacing(np.abs(p[kind == 5])).astype(np.float32)              # a few ulps off the surface
    out[kind == 5] = p[kind == 5] + nn[kind == 5] * ulp * rng.integers(-4, 5, (kind == 5).sum())[:, None].astype(np.float32)
    return np.ascontiguousarray(out, np.float32)


def _compare(engine, oracle, comp, pts, name):
    engine.load_renderer(comp)
    lib = _lib.load()
    got = np.zeros(len(pts), np.int32)
    _lib.check(lib.pt_debug_region_at(engine.ctx, pts.ctypes.data, len(pts), got.ctypes.data))
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    want = np.array([ors.region_at(p) for p in pts], np.int32)
    bad = np.nonzero(got != want)[0]
    _say(f"[gpu] {name}: {len(pts)} points, {np.mean(want >= 0):.3f} inside a region, {len(bad)} differ")
    assert len(bad) == 0, (bad[:5], got[bad[:5]], want[bad[:5]], pts[bad[:5]])
    return want


def test_probe_map_equals_the_walk_on_the_icosphere_cloud(oracle):
    rng = np.random.default_rng(17)
    sc = scenes.icosphere_cloud()
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, 3), 16, 16)
    tris = comp.scene.tri_vertices
    pts = np.concatenate([_surface_points(tris, 1 << 17, rng),
                          rng.uniform(0.0, 100.0, size=(1 << 14, 3)).astype(np.float32)])
    with Engine(0) as e:
        want = _compare(e, oracle, comp, pts, "icosphere cloud")
    assert 0.1 < np.mean(want >= 0) < 0.9


def test_probe_map_equals_the_walk_on_interpenetrating_spheres(oracle):
    rng = np.random.default_rng(23)
    sc = scenes.icosphere_cloud(n_spheres=300, seed=3, box=12.0)
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, 3), 16, 16)
    pts = np.concatenate([_surface_points(comp.scene.tri_vertices, 1 << 16, rng),
                          rng.uniform(0.0, 12.0, size=(1 << 14, 3)).astype(np.float32)])
    with Engine(0) as e:
        want = _compare(e, oracle, comp, pts, "dense cloud")
    assert np.mean(want >= 0) > 0.2


def test_probe_map_equals_the_walk_through_placements(oracle):
    rng = np.random.default_rng(29)
    sc = scenes.cornell_glass_instances(n=6)
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, 3), 16, 16)
    # world-space copies of the placed prototype triangles
    batch = [o for o in sc["objects"] if o["type"] == "instances"][0]
    proto = np.asarray(batch["triangles"], np.float64).reshape(-1, 3, 3)
    world = []
    for pl in batch["placements"]:
        world.append(scenes.transformed_triangles(proto.reshape(-1, 9), pl["rotation"], pl["translation"], pl["scale"]))
    tris = np.concatenate(world).astype(np.float32).reshape(-1, 9)
    pts = np.concatenate([_surface_points(tris, 1 << 15, rng),
                          rng.uniform(-1.0, 1.0, size=(1 << 13, 3)).astype(np.float32)])
    with Engine(0) as e:
        want = _compare(e, oracle, comp, pts, "glass placements")
    assert np.mean(want >= 0) > 0.05
