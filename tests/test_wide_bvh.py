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


def test_leaf_caps_and_node_budget():
    """cwbvh.test.ts's caps: a leaf slot holds 1..3 triangles, a node's leaf triangles fit the
    24 addressable offsets, every node has a child (the SAH collapse never emits an empty node)."""
    for n, seed in ((9, 7), (300, 8), (5000, 9)):
        nodes, tris, _ = wide_bvh_ref.build(_soup(n, seed))
        for k in range(len(nodes)):
            ch = wide_bvh_ref.decode(nodes[k])
            assert len(ch) > 0
            for c in ch:
                if not c["internal"]:
                    cnt = len(c["tris"])
                    assert 1 <= cnt <= 3 and (c["meta"] & 31) + cnt <= 24


def test_adversarial_extents():
    """cwbvh.test.ts 'adversarial extents': zero-width axes, tiny and huge triangles; decoded
    boxes still contain every triangle and every triangle appears once."""
    rng = np.random.default_rng(11)
    flat = _soup(600, 12).reshape(-1, 3, 3)
    flat[:, :, 2] = 0.0                                           # every triangle in z = 0
    check(flat.reshape(-1, 9))
    line = np.zeros((200, 3, 3), np.float32)
    line[:, :, 0] = rng.uniform(-5, 5, size=(200, 3))             # y = z = 0: a segment cloud
    check(line.reshape(-1, 9))
    tiny = (rng.normal(size=(300, 3, 3)) * 1e-12 + rng.uniform(-1, 1, size=(300, 1, 3))).astype(np.float32)
    check(tiny.reshape(-1, 9))
    huge = (rng.normal(size=(300, 3, 3)) * 1e18).astype(np.float32)
    check(huge.reshape(-1, 9))


def test_depth_stays_far_below_the_stack():
    """cwbvh.test.ts: a 50k-item cloud sits far below the walk's stack (depth < 12)."""
    depth = check(_soup(50000, 13))
    assert depth < 12
