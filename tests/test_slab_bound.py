"""The wide-BVH node test's rounding bound (csrc/traversal.h trace_step), restated in fp32 C in
the oracle (or_slab_probe) and checked against the exact slab distances of the same quantized
boxes: the fp32 entry distance is never above the exact one and the fp32 exit distance never
below it, for random boxes and rays and for adversarial ones (origins on and one ulp off the
planes, far origins, axis-aligned and near-axis directions, degenerate and wide exponents), so
a child is never culled where the exact test of its box would enter it. The slack is also
bounded: at most 14 u (|oq| + 255 |dq|) per side. With the bound set to zero the same cases
violate the containment on about half of the rays (checked once on the CPU), so the check is
not vacuous."""
import numpy as np

import oracle_lib

U = 2.0 ** -24


def exact_interval(rays, boxes):
    """Entry / exit distances of the exact slab test in float64 (planes p + q 2^e are exact in
    f64 for these ranges; one rounding per division), with the kernel's direction guard."""
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    d = np.where(np.abs(d) < 1e-20, np.copysign(1e-20, d), d)
    p, e = boxes[:, 0:3].astype(np.float64), boxes[:, 3:6].astype(np.float64)
    lo = p + boxes[:, 6:9].astype(np.float64) * np.exp2(e)
    hi = p + boxes[:, 9:12].astype(np.float64) * np.exp2(e)
    t0, t1 = (lo - o) / d, (hi - o) / d
    te = np.maximum(np.minimum(t0, t1).max(axis=1), 0.0)
    tf = np.maximum(t0, t1).min(axis=1)
    scale = ((np.abs(p - o) + 255.0 * np.exp2(e)) / np.abs(d)).max(axis=1)
    return te, tf, scale


def cases(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1000, 1000, size=(n, 3)) * (10.0 ** rng.integers(-3, 1, size=(n, 1)))
    e = rng.integers(-30, 8, size=(n, 3)).astype(np.float64)
    q = np.sort(rng.integers(0, 256, size=(n, 2, 3)), axis=1).astype(np.float64)
    boxes = np.concatenate([p, e, q[:, 0], q[:, 1]], axis=1).astype(np.float32)
    lo = p + q[:, 0] * np.exp2(e)
    hi = p + q[:, 1] * np.exp2(e)
    kind = rng.integers(0, 4, size=n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ax = rng.integers(0, 3, size=n)
    axis_aligned = rng.random(n) < 0.15
    d[axis_aligned] = 0.0
    d[axis_aligned, ax[axis_aligned]] = rng.choice([-1.0, 1.0], size=axis_aligned.sum())
    near_axis = (~axis_aligned) & (rng.random(n) < 0.15)
    d[near_axis, ax[near_axis]] *= rng.choice([1e-9, 1e-25, 0.0, -0.0], size=near_axis.sum())
    target = lo + rng.random((n, 3)) * (hi - lo)
    dist = np.where(kind[:, None] == 0, rng.uniform(0, 5, size=(n, 1)) * (hi - lo).max(axis=1, keepdims=True),
                    rng.uniform(1.0, 5000.0, size=(n, 1)))
    o = target - d * dist
    # kind 2: the origin on a plane (snapped), kind 3: one ulp off it
    on = kind >= 2
    which = rng.integers(0, 2, size=n)
    plane = np.where(which[:, None] == 0, lo, hi)
    o[on, ax[on]] = plane[on, ax[on]]
    o = o.astype(np.float32)
    off = kind == 3
    o[off, ax[off]] = np.nextafter(o[off, ax[off]], np.where(rng.random(off.sum()) < 0.5, -np.inf, np.inf).astype(np.float32))
    rays = np.concatenate([o, d.astype(np.float32)], axis=1).astype(np.float32)
    return np.ascontiguousarray(rays), np.ascontiguousarray(boxes)


def test_slab_bound_is_conservative_and_tight(oracle):
    total = 0
    for seed in range(4):
        rays, boxes = cases(200000, seed)
        out = np.zeros((len(rays), 2), np.float32)
        oracle.or_slab_probe(rays.ctypes.data, boxes.ctypes.data, len(rays), out.ctypes.data)
        te, tf, scale = exact_interval(rays, boxes)
        k_te, k_tf = out[:, 0].astype(np.float64), out[:, 1].astype(np.float64)
        fin = np.isfinite(te) & np.isfinite(tf) & np.isfinite(k_te) & np.isfinite(k_tf)
        assert fin.mean() > 0.99
        assert (k_te[fin] <= te[fin]).all(), np.nonzero(k_te[fin] > te[fin])[0][:5]
        assert (k_tf[fin] >= tf[fin]).all(), np.nonzero(k_tf[fin] < tf[fin])[0][:5]
        # every box the exact test enters is entered by the fp32 test
        enter = fin & (te <= tf)
        assert (k_te[enter] <= k_tf[enter]).all()
        # the slack is a few ulps of the plane distances, not more
        assert ((te - k_te)[enter & (te > 0)] <= 14 * U * scale[enter & (te > 0)] + 1e-29).all()
        assert ((k_tf - tf)[fin] <= 14 * U * scale[fin] + 1e-29).all()
        total += int(enter.sum())
    assert total > 100000
