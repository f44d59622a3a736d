"""Structure of the wide BVH the library builds (host code, no GPU): every triangle appears
exactly once, indices stay in range, decoded child boxes contain their geometry, and the
leaf records reproduce e1 = b - a, e2 = c - a bit for bit."""
import numpy as np

import wide_bvh_ref
from pathtracerglsl_amd import scenes


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, size=(n, 1, 3))
    return (c + rng.normal(scale=0.5, size=(n, 3, 3))).reshape(n, 9).astype(np.float32)


def check(tv):
    nodes, tris, depth = wide_bvh_ref.build(tv)
    ids = tris[:, 3].view(np.uint32)
    assert sorted(ids.tolist()) == list(range(len(tv)))
    a = tv[ids, 0:3]
    assert np.array_equal(tris[:, 0:3], a)
    assert np.array_equal(tris[:, 4:7], tv[ids, 3:6] - a)
    assert np.array_equal(tris[:, 8:11], tv[ids, 6:9] - a)
    seen_nodes = {0}
    stack = [0]
    covered = 0
    while stack:
        k = stack.pop()
        for ch in wide_bvh_ref.decode(nodes[k]):
            if ch["internal"]:
                assert 0 < ch["child"] < len(nodes) and ch["child"] not in seen_nodes
                seen_nodes.add(ch["child"])
                stack.append(ch["child"])
            else:
                for t in ch["tris"]:
                    assert 0 <= t < len(tris)
                    v = tv[ids[t]].reshape(3, 3)
                    assert (v >= ch["lo"]).all() and (v <= ch["hi"]).all()
                    covered += 1
    assert covered == len(tv) and len(seen_nodes) == len(nodes)
    return depth


def test_small_soups():
    for n, seed in ((1, 0), (2, 1), (3, 2), (7, 3), (25, 4), (500, 5), (4000, 6)):
        check(_soup(n, seed))


def test_degenerate_and_coincident():
    tv = np.zeros((40, 9), np.float32)        # 40 identical degenerate triangles
    check(tv)
    tv = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (33, 1))
    check(tv)


def test_icosphere_cloud_structure():
    sc = scenes.icosphere_cloud(n_spheres=200)
    tv = np.concatenate([o["triangles"].reshape(-1, 9) for o in sc["objects"]]).astype(np.float32)
    depth = check(tv)
    assert depth <= 16
