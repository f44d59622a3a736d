"""Synthetic SceneDescriptions for the benchmark workloads.

Each builder returns a SceneDescription dict in the shape the reference's authoring layer
produces (objects with materials, lights, environment, camera; src/compiler/types.ts:561-721),
with meshes already flattened to world-space triangle soups. All randomness is seeded.
"""
import math

import numpy as np

# ---------------------------------------------------------------- Cornell box (workload C0)


def _quad_tris(c, e1, e2):
    """Two triangles covering corner c + s e1 + t e2, wound so cross(e1, e2) is the normal."""
    c, e1, e2 = (np.asarray(v, np.float64) for v in (c, e1, e2))
    a, b, d, e = c, c + e1, c + e1 + e2, c + e2
    return [[a, b, d], [a, d, e]]


def _box_tris(lo, hi):
    """Closed axis-aligned box, outward winding, 12 triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    tris = []
    tris += _quad_tris((x0, y0, z1), (x1 - x0, 0, 0), (0, y1 - y0, 0))      # +z
    tris += _quad_tris((x1, y0, z0), (x0 - x1, 0, 0), (0, y1 - y0, 0))      # -z
    tris += _quad_tris((x1, y0, z1), (0, 0, z0 - z1), (0, y1 - y0, 0))      # +x
    tris += _quad_tris((x0, y0, z0), (0, 0, z1 - z0), (0, y1 - y0, 0))      # -x
    tris += _quad_tris((x0, y1, z1), (x1 - x0, 0, 0), (0, 0, z0 - z1))      # +y
    tris += _quad_tris((x0, y0, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0))      # -y
    return np.asarray(tris, np.float64)


def cornell_box(light_radiance=(17.0, 12.0, 4.0)):
    """Open-front box [-1,1]^3 (10 triangles) + two axis-aligned inner boxes (24 triangles),
    Lambert white/red/green, a 0.5 x 0.5 ceiling quad light at y = 0.999 facing -y."""
    white = _quad_tris((-1, -1, 1), (2, 0, 0), (0, 0, -2))        # floor
    white += _quad_tris((-1, 1, -1), (2, 0, 0), (0, 0, 2))        # ceiling
    white += _quad_tris((-1, -1, -1), (2, 0, 0), (0, 2, 0))       # back
    red = _quad_tris((-1, -1, 1), (0, 0, -2), (0, 2, 0))          # left
    green = _quad_tris((1, -1, -1), (0, 0, 2), (0, 2, 0))         # right
    return {
        "id": "cornell",
        "materials": {
            "white": {"model": "lambert", "albedo": [0.73, 0.73, 0.73]},
            "red": {"model": "lambert", "albedo": [0.63, 0.065, 0.05]},
            "green": {"model": "lambert", "albedo": [0.14, 0.45, 0.09]},
        },
        "objects": [
            {"type": "mesh", "triangles": np.asarray(white), "material": "white", "closed": False},
            {"type": "mesh", "triangles": np.asarray(red), "material": "red", "closed": False},
            {"type": "mesh", "triangles": np.asarray(green), "material": "green", "closed": False},
            {"type": "mesh", "triangles": _box_tris((0.05, -1.0, 0.0), (0.65, -0.4, 0.6)),
             "material": "white", "closed": True},
            {"type": "mesh", "triangles": _box_tris((-0.65, -1.0, -0.6), (-0.05, 0.2, 0.0)),
             "material": "white", "closed": True},
        ],
        "lights": [
            {"type": "quad", "corner": [-0.25, 0.999, -0.25], "edge1": [0.5, 0.0, 0.0],
             "edge2": [0.0, 0.0, 0.5], "radiance": list(light_radiance)},
        ],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 3.5], "target": [0.0, 0.0, 0.0],
                   "fov": math.radians(40.0)},
    }


def two_light():
    """The reference's two-light witness (tests/witnesses/scenes/twoLightScene.ts): two diffuse
    spheres on a floor before a back wall, lit only by a warm key and a cool fill point light of
    unequal power, posed at (0, 1.2, 4) -> (0, 0.5, 0) with fov 0.9 (tests/witnesses/index.ts:
    178-191). The reference's planes (floor y = 0, wall z = -3) are 40 x 40 quads here."""
    return {
        "id": "two-light",
        "materials": {
            "floor": {"model": "lambert", "albedo": [0.6, 0.6, 0.6]},
            "ivory": {"model": "lambert", "albedo": [0.82, 0.78, 0.68]},
        },
        "objects": [
            {"type": "quad", "corner": [-20.0, 0.0, 20.0], "edge1": [40.0, 0.0, 0.0], "edge2": [0.0, 0.0, -40.0],
             "material": "floor"},
            {"type": "quad", "corner": [-20.0, 0.0, -3.0], "edge1": [40.0, 0.0, 0.0], "edge2": [0.0, 40.0, 0.0],
             "material": "floor"},
            {"type": "sphere", "center": [-0.75, 0.5, 0.0], "radius": 0.5, "material": "ivory"},
            {"type": "sphere", "center": [0.75, 0.5, 0.0], "radius": 0.5, "material": "ivory"},
        ],
        "lights": [
            {"type": "point", "position": [-2.5, 3.0, 1.5], "intensity": [26.0, 14.3, 6.5]},
            {"type": "point", "position": [2.5, 2.5, 1.5], "intensity": [4.9, 7.7, 14.0]},
        ],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 1.2, 4.0], "target": [0.0, 0.5, 0.0], "fov": 0.9},
    }


def many_lights(n=64, seed=3, rough_glass=False):
    """A many-lights scene in the spirit of the reference's hundred-spheres witness
    (tests/witnesses/index.ts:192-211), with n small one-sided quad lights of assorted colours
    and powers scattered over a 12 x 12 floor at heights 0.4-2.5, facing down or tilted, some
    below a raised slab, around two diffuse spheres (a rough glass sphere instead of one with
    rough_glass, whose two-sided light queries disarm the horizon cull). Posed from a corner,
    as the witness is, so that selection by distance matters."""
    rng = np.random.default_rng(seed)
    lights = []
    for k in range(n):
        x, z = rng.uniform(-6.0, 6.0, 2)
        y = rng.uniform(0.4, 2.5)
        size = rng.uniform(0.08, 0.3)
        tilt = rng.uniform(-0.6, 0.6)
        e1 = [size, 0.0, 0.0]
        e2 = [0.0, size * tilt, -size] if k % 3 else [0.0, -size * tilt, size]
        col = rng.uniform(0.2, 1.0, 3) * rng.uniform(5.0, 60.0)
        lights.append({"type": "quad", "corner": [float(x), float(y), float(z)], "edge1": e1,
                       "edge2": [float(v) for v in e2], "radiance": [float(c) for c in col]})
    mats = {"floor": {"model": "lambert", "albedo": [0.55, 0.55, 0.55]},
            "clay": {"model": "lambert", "albedo": [0.7, 0.45, 0.3]},
            "slab": {"model": "lambert", "albedo": [0.3, 0.5, 0.7]}}
    objs = [{"type": "quad", "corner": [-8.0, 0.0, 8.0], "edge1": [16.0, 0.0, 0.0], "edge2": [0.0, 0.0, -16.0],
             "material": "floor"},
            {"type": "quad", "corner": [-2.0, 1.2, 3.0], "edge1": [4.0, 0.0, 0.0], "edge2": [0.0, 0.0, -2.0],
             "material": "slab"},
            {"type": "sphere", "center": [-1.0, 0.6, 0.0], "radius": 0.6, "material": "clay"}]
    if rough_glass:
        mats["glass"] = {"model": "rough_dielectric", "ior": 1.5, "roughness": 0.3}
        objs.append({"type": "sphere", "center": [1.2, 0.5, -0.5], "radius": 0.5, "material": "glass"})
    else:
        objs.append({"type": "sphere", "center": [1.2, 0.5, -0.5], "radius": 0.5, "material": "clay"})
    return {"id": "many-lights", "materials": mats, "objects": objs, "lights": lights,
            "environment": {"type": "none"},
            "camera": {"type": "pinhole", "position": [-7.5, 1.8, -7.5], "target": [-4.0, 0.6, -4.0], "fov": 0.9}}


def cornell_multi_light(point=True):
    """The Cornell box lit by several finite lights of unequal power: the ceiling quad (17, 12, 4),
    a smaller, cooler quad on the back wall facing +z, a dim warm quad low on the left wall
    facing +x and (point=True) a point light under the ceiling: the N-light selection CDF, its
    cdf_rescale and the per-light selection pdf in the MIS weight all act."""
    sc = cornell_box()
    sc["id"] = "cornell-multi-light" + ("" if point else "-quads")
    sc["lights"] = sc["lights"] + [
        {"type": "quad", "corner": [0.3, 0.3, -0.999], "edge1": [0.3, 0.0, 0.0], "edge2": [0.0, 0.3, 0.0],
         "radiance": [3.0, 6.0, 12.0]},
        {"type": "quad", "corner": [-0.999, -0.9, 0.2], "edge1": [0.0, 0.2, 0.0], "edge2": [0.0, 0.0, 0.4],
         "radiance": [6.0, 3.0, 1.0]},
    ]
    if point:
        sc["lights"].append({"type": "point", "position": [0.4, 0.7, 0.5], "intensity": [0.6, 0.5, 0.4]})
    return sc


# ---------------------------------------------------------- icosphere cloud (workload C1)

_ICO_T = (1.0 + math.sqrt(5.0)) / 2.0
_ICO_V = np.array([[-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
                   [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
                   [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]], np.float64)
_ICO_F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                   [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                   [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)


def unit_icosphere(detail=4):
    """Indexed unit icosphere, three.js-style subdivision: every icosahedron face is split
    into (detail+1)^2 triangles and projected to the sphere (20 (d+1)^2 faces). Shared
    vertices are deduplicated so the mesh is watertight; winding is outward."""
    n = detail + 1
    verts, keys, faces = [], {}, []

    def vid(p):
        p = p / np.linalg.norm(p)
        k = tuple(np.round(p, 9))
        if k not in keys:
            keys[k] = len(verts)
            verts.append(p)
        return keys[k]

    for fa, fb, fc in _ICO_F:
        a, b, c = _ICO_V[fa], _ICO_V[fb], _ICO_V[fc]
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                grid[i, j] = vid(a + (b - a) * (i / n) + (c - a) * (j / n))
        for i in range(n):
            for j in range(n - i):
                faces.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j + 1 < n:
                    faces.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    V = np.asarray(verts)
    F = np.asarray(faces, np.int64)
    # outward winding check: (b-a)x(c-a) . a > 0
    g = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    if np.mean(np.einsum("ij,ij->i", g, V[F[:, 0]])) < 0:
        F = F[:, [0, 2, 1]]
    return V, F


def _value_noise(p, seed):
    """Seeded 3D value noise in [0, 1): hashed lattice values, smoothstep trilinear blend."""
    p = np.asarray(p, np.float64)
    i0 = np.floor(p).astype(np.int64)
    f = p - i0
    w = f * f * (3.0 - 2.0 * f)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return (x & 0xFFFFFF).astype(np.float64) / float(1 << 24)

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = w[..., 0] if dx else 1.0 - w[..., 0]
                wy = w[..., 1] if dy else 1.0 - w[..., 1]
                wz = w[..., 2] if dz else 1.0 - w[..., 2]
                out = out + wx * wy * wz * h(i0[..., 0] + dx, i0[..., 1] + dy, i0[..., 2] + dz)
    return out


def icosphere_cloud(n_spheres=2000, detail=4, seed=7, box=100.0, displacement=0.15):
    """2000 noise-displaced detail-4 icospheres (500 triangles each, 1.0 M total), centres
    uniform in [0, box]^3, radii lognormal(0, 0.5); 60 % GGX (roughness U[0.05, 0.8]),
    20 % smooth dielectric eta 1.5, 20 % Lambert (albedo U[0.1, 0.9]); procedural sky."""
    rng = np.random.default_rng(seed)
    V, F = unit_icosphere(detail)
    centres = rng.uniform(0.0, box, size=(n_spheres, 3))
    radii = rng.lognormal(0.0, 0.5, size=n_spheres)
    kinds = rng.choice(3, size=n_spheres, p=[0.6, 0.2, 0.2])
    materials, objects = {}, []
    for s in range(n_spheres):
        noise = _value_noise(V * 2.5 + rng.uniform(0, 64, size=3), seed + s)
        r = radii[s] * (1.0 + displacement * (2.0 * noise - 1.0))
        P = centres[s] + V * r[:, None]
        name = f"m{s}"
        if kinds[s] == 0:
            materials[name] = {"model": "ggx", "f0": rng.uniform(0.3, 0.95, size=3).tolist(),
                               "roughness": float(rng.uniform(0.05, 0.8))}
        elif kinds[s] == 1:
            materials[name] = {"model": "dielectric", "ior": 1.5}
        else:
            materials[name] = {"model": "lambert", "albedo": [float(rng.uniform(0.1, 0.9))] * 3}
        objects.append({"type": "mesh", "triangles": P[F], "material": name, "closed": True})
    return {
        "id": f"icosphere-cloud-{n_spheres}",
        "materials": materials,
        "objects": objects,
        "lights": [],
        "environment": {"type": "procedural", "intensity": 1.0},
        "camera": {"type": "pinhole", "position": [box * 0.5, box * 0.5, -0.05 * box],
                   "target": [box * 0.5, box * 0.5, box * 0.5], "fov": math.radians(60.0)},
    }


def random_rays(n, seed=1, box=100.0):
    """Seeded hit-parity ray set: origins uniform in the box, directions uniform on S^2."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.0, box, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


# ----------------------------------------------------------------- witness scenes (tests)

def mesh_furnace():
    """F-BOX through the triangle engine (tests/witnesses/scenes/meshWitness.ts:64-81): a
    cube wound inward, Lambert albedo 0.5 + emission 0.2, camera at the centre. Every path
    hits a wall at every bounce, so each pixel is exactly sum_{n<=N} 0.2 * 0.5^n -> 0.4."""
    cube = _box_tris((-1, -1, -1), (1, 1, 1))[:, [0, 2, 1]]      # inward winding
    return {
        "id": "mesh-furnace",
        "materials": {"furnace": {"model": "lambert", "albedo": [0.5, 0.5, 0.5], "emission": [0.2, 0.2, 0.2]}},
        "objects": [{"type": "mesh", "triangles": cube, "material": "furnace", "closed": False}],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 0.0], "target": [0.0, 0.0, -1.0], "fov": 0.8},
    }


def eta_witness():
    """F-ETA (tests/witnesses/scenes/dielectricWitness.ts:15-53): an emitter under a flat water
    surface (eta 1.33), camera in air looking down. Centre pixel = (1 - R0) / 1.33^2 = 0.5540;
    without the eta^2 factor it would read 0.980. The water half-space is a large closed box,
    the emitter a quad at y = -1 facing up."""
    water = _box_tris((-60.0, -3.0, -60.0), (60.0, 0.0, 60.0))
    return {
        "id": "eta",
        "materials": {"water": {"model": "dielectric", "ior": 1.33},
                      "glow": {"model": "lambert", "albedo": [0.0, 0.0, 0.0], "emission": [1.0, 1.0, 1.0]}},
        "objects": [
            {"type": "mesh", "triangles": water, "material": "water", "closed": True},
            {"type": "quad", "corner": [-50.0, -1.0, 50.0], "edge1": [100.0, 0.0, 0.0], "edge2": [0.0, 0.0, -100.0],
             "material": "glow"},
        ],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 1.0, 0.05], "target": [0.0, -1.0, 0.0], "fov": 0.6},
    }


def slab_witness():
    """F-SLAB (tests/witnesses/scenes/mediaWitness.ts:15-53): an emissive screen behind an
    absorbing ink slab z in [-1, 0] bounded by null interfaces (model 'none'); the camera looks
    straight through, so the centre pixel is exactly exp(-sigma_a * 1) = (0.36788, 0.13534,
    0.01832). The plane screen is a large emissive quad facing +z."""
    return {
        "id": "slab",
        "materials": {"screen": {"model": "lambert", "albedo": [0.0, 0.0, 0.0], "emission": [1.0, 1.0, 1.0]},
                      "ink": {"model": "none", "medium": {"sigma_a": [1.0, 2.0, 4.0]}}},
        "objects": [
            {"type": "quad", "corner": [-50.0, -50.0, -2.0], "edge1": [100.0, 0.0, 0.0], "edge2": [0.0, 100.0, 0.0],
             "material": "screen"},
            {"type": "mesh", "triangles": _box_tris((-4.0, -4.0, -1.0), (4.0, 4.0, 0.0)), "material": "ink", "closed": True},
        ],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 2.0], "target": [0.0, 0.0, -2.0], "fov": 0.6},
    }


def furnace_scatter():
    """F-BOX-M (mediaWitness.ts:54-100): the F-BOX furnace filled with a non-absorbing,
    chromatic, anisotropic ambient haze. A non-absorbing medium keeps the equilibrium field,
    so every channel is still exactly 0.4 in expectation."""
    sc = mesh_furnace()
    sc["id"] = "furnace-scatter"
    sc["materials"]["haze"] = {"model": "none", "medium": {"sigma_a": 0.0, "sigma_s": [0.5, 1.0, 2.0], "phase_g": 0.7}}
    sc["ambient_medium"] = "haze"
    sc["camera"]["fov"] = 1.0
    return sc


def rayleigh_furnace(sigma_s=(0.1, 0.2, 0.4)):
    """spectral-rayleigh-furnace (tests/witnesses/scenes/spectralMedia.ts:65-73): the F-BOX-M haze
    with the Rayleigh phase (parameter-free, volume_scattering/rayleigh). The phase law changes
    without changing the furnace's equilibrium: every channel is 0.4 in expectation."""
    sc = furnace_scatter()
    sc["id"] = "rayleigh-furnace"
    sc["materials"]["haze"] = {"model": "none", "medium": {"model": "rayleigh", "sigma_a": 0.0, "sigma_s": list(sigma_s)}}
    return sc


GAUSSIAN_BAND = {"gaussian": {"center": 610.0, "width": 25.0, "peak": 0.85, "floor": 0.04}}


def gaussian_plane(band=None):
    """spectral-gaussian-plane (tests/witnesses/scenes/spectralGaussian.ts): a Lambert plane with a
    Gaussian reflectance band under a unit white constant sky sampled as a light, seen head-on. The
    one-bounce pixel is the band's RGB stand-in: in a spectral program the integral of sensor x band
    x D65, in an RGB program the stand-in itself."""
    return {
        "id": "spectral-gaussian-plane",
        "materials": {"plane": {"model": "lambert", "albedo": band or GAUSSIAN_BAND}},
        "objects": [{"type": "quad", "corner": [-50.0, -50.0, -1.0], "edge1": [100.0, 0.0, 0.0],
                     "edge2": [0.0, 100.0, 0.0], "material": "plane"}],
        "lights": [],
        "environment": {"type": "constant", "color": [1.0, 1.0, 1.0], "intensity": 1.0, "sample_as_light": True},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 0.0], "target": [0.0, 0.0, -1.0], "fov": 0.8},
    }


def _ground(y=0.0, half=200.0):
    """The infinite plane y = const of the reference's scenes, as a 2 half x 2 half quad facing +y."""
    return {"type": "quad", "corner": [-half, y, half], "edge1": [2.0 * half, 0.0, 0.0],
            "edge2": [0.0, 0.0, -2.0 * half]}


def proc_sky(glsl=None, rotation=0.0, chart="equirect", compensated=False):
    """proc-sky (tests/witnesses/scenes/envScenes.ts:10-40, 172-187): a Lambert ground (albedo
    0.4), a clay sphere and a glass sphere under a procedural sky (the proc-sky body unless
    `glsl` names another), posed from (0, 1.4, 5) at (0, 0.9, 0), fov 0.9."""
    env = {"type": "procedural", "intensity": 1.0, "rotation": rotation, "chart": chart, "compensated": compensated}
    if glsl is not None:
        env["glsl"] = {"kind": "glsl", "source": glsl}
    return {
        "id": "proc-sky",
        "materials": {"ground": {"model": "lambert", "albedo": [0.4, 0.4, 0.4]},
                      "clay": {"model": "lambert", "albedo": [0.7, 0.45, 0.3]},
                      "glass": {"model": "dielectric", "ior": 1.5}},
        "objects": [{**_ground(), "material": "ground"},
                    {"type": "sphere", "center": [0.0, 1.0, 0.0], "radius": 1.0, "material": "clay"},
                    {"type": "sphere", "center": [2.2, 0.7, -0.5], "radius": 0.7, "material": "glass"}],
        "lights": [],
        "environment": env,
        "camera": {"type": "pinhole", "position": [0.0, 1.4, 5.0], "target": [0.0, 0.9, 0.0], "fov": 0.9},
    }


# proc-sky-rotated's body (tests/witnesses/scenes/estimatorAgreementWitness.ts:61-72): a 10-degree
# sun of radiance 40 over a sky of 0.02
BIG_SUN_SKY = "vec3(0.02) + vec3(40.0) * smoothstep(0.985, 0.99, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))"


def proc_sky_rotated(chart="equirect"):
    """proc-sky-rotated (estimatorAgreementWitness.ts:61-72): proc-sky under the big-sun body
    turned by -3 rad, which puts the sun in the band of table columns whose u = phi / 2 pi + 1/2 +
    r / 2 pi leaves [0, 1): the environment pdf must wrap u as the sampler does, so pt-nee and
    pt-mis converge to one image."""
    sc = proc_sky(BIG_SUN_SKY, rotation=-3.0, chart=chart)
    sc["id"] = "proc-sky-rotated"
    return sc


def sky_plane(glsl, scene_id, chart="equirect", compensated=False, table_size=(4, 2)):
    """The spectral-skies plane (tests/witnesses/scenes/spectralSkies.ts:12-16): a Lambert plane
    y = 0 (albedo 0.5) seen from (0, 1, 0) toward (0, 0, -0.1), fov 0.6, under a procedural sky
    with a tiny importance table."""
    return {
        "id": scene_id,
        "materials": {"floor": {"model": "lambert", "albedo": [0.5, 0.5, 0.5]}},
        "objects": [{**_ground(), "material": "floor"}],
        "lights": [],
        "environment": {"type": "procedural", "intensity": 1.0, "chart": chart, "compensated": compensated,
                        "table_size": list(table_size), "glsl": {"kind": "glsl", "source": glsl}},
        "camera": {"type": "pinhole", "position": [0.0, 1.0, 0.0], "target": [0.0, 0.0, -0.1], "fov": 0.6},
    }


# spectral-image-constant's picture (tests/witnesses/scenes/spectralSkies.ts:18-23): 16 x 16 texels of
# exactly this linear RGB (the RGBE bytes 32, 64, 128, 129)
IMAGE_CONSTANT_RGB = (0.25, 0.5, 1.0)


def sun_picture(w=64, h=32, seed=5):
    """A synthetic equirect sky picture: a vertical gradient, noise, and a hot 3 x 2 sun patch."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    pic = np.empty((h, w, 3), np.float32)
    pic[:] = (0.1 + 0.5 * (1 - v))[:, None, None] * np.array([0.6, 0.8, 1.0], np.float32)
    pic *= rng.uniform(0.8, 1.2, size=(h, w, 1)).astype(np.float32)
    pic[h // 4: h // 4 + 2, (2 * w) // 3: (2 * w) // 3 + 3] = [60.0, 50.0, 35.0]
    return pic


def image_plane(image=None, intensity=2.0, rotation=0.73, chart="equirect", compensated=False, table_n=256):
    """spectral-image-constant (spectralSkies.ts:53-66): the Lambert plane (albedo 0.5) of the
    spectral-skies suite under an image environment, by default the constant picture of
    IMAGE_CONSTANT_RGB at intensity 2 and rotation 0.73, so the plane's radiance is that RGB."""
    if image is None:
        image = np.broadcast_to(np.array(IMAGE_CONSTANT_RGB, np.float32), (16, 16, 3)).copy()
    sc = sky_plane("vec3(1.0)", "spectral-image-constant", chart, compensated)
    sc["environment"] = {"type": "image", "image": image, "intensity": intensity, "rotation": rotation,
                         "chart": chart, "compensated": compensated, "table_size": [2 * table_n, table_n]}
    return sc


def checker_plane(corner=(-0.5, 0.0, 0.5), size=3.0, uv_scale=3.0, albedo_a=(0.9, 0.6, 0.3), albedo_b=(0.1, 0.2, 0.4)):
    """The spectral-skies plane as a size x size checker quad (materials/checker) under a white sky: in
    pt with one bounce each pixel is the albedo of the cell its hit lies in, exactly. The default
    corner puts every visible hit in cell (0, 0) (albedo_a); corner (-1.5, 0, 0.5) in cell (1, 0)
    (albedo_b)."""
    sc = sky_plane("vec3(1.0)", "checker-plane")
    sc["materials"] = {"floor": {"model": "checker", "albedo_a": list(albedo_a), "albedo_b": list(albedo_b),
                                 "uv_scale": uv_scale}}
    sc["objects"] = [{"type": "quad", "corner": list(corner), "edge1": [size, 0.0, 0.0], "edge2": [0.0, 0.0, -size],
                      "material": "floor"}]
    sc["environment"] = {"type": "constant", "color": [1.0, 1.0, 1.0], "intensity": 1.0, "sample_as_light": True}
    return sc


def checker_room(detail=2):
    """A checker floor quad, a checker sphere and a checker mesh (the planar placeholder chart) in
    the Cornell box's light."""
    sc = cornell_box()
    sc["materials"]["check"] = {"model": "checker", "albedo_a": [0.8, 0.8, 0.2], "albedo_b": [0.1, 0.3, 0.7],
                                "uv_scale": 6.0}
    sc["objects"].append({"type": "sphere", "center": [-0.45, -0.7, -0.2], "radius": 0.3, "material": "check"})
    v, f = unit_icosphere(detail)
    tris = (v[f] * 0.25 + np.array([0.4, -0.7, 0.2])).reshape(-1, 9)
    sc["objects"].append({"type": "mesh", "triangles": tris.tolist(), "material": "check"})
    sc["objects"].append({"type": "quad", "corner": [-0.9, -0.99, 0.9], "edge1": [1.8, 0.0, 0.0], "edge2": [0.0, 0.0, -1.8],
                          "material": "check"})
    return sc


def sky_missed_cap(chart="equirect"):
    """sky-missed-cap (spectralSkies.ts:38-46): the sky is 1 inside the polar cap dir.y >= 0.9 and
    0 elsewhere, and a 4 x 2 table (2 x 2 octahedral) has no texel centre in the cap, so the table
    holds no mass and the sampler draws the uniform sphere alone. The plane's radiance is
    rho (1 - cos^2 theta_cap) = 0.5 (1 - 0.81) = 0.095 under NEE."""
    return sky_plane("vec3(step(0.9, dir.y))", "sky-missed-cap", chart)


def spectral_comp_empty(chart="equirect"):
    """spectral-comp-empty (spectralSkies.ts:47-52): a white sky's MIS-compensated table is empty,
    so environment NEE draws nothing and the BSDF's misses take full MIS weight: the plane's
    radiance is rho L = 0.5."""
    return sky_plane("vec3(1.0)", "spectral-comp-empty", chart, compensated=True)


def cornell_glass_medium(light_radiance=(12.0, 12.0, 12.0)):
    """The spectral-media workload's geometry in RGB: the Cornell box with a radius-0.5 smooth
    dielectric sphere (eta 1.5) at the centre, filled with a chromatic homogeneous medium
    sigma_t = (0.8, 1.2, 2.0), albedo 0.9, HG g = 0.7 (analytic channel-MIS distance sampling;
    the glass is opaque to shadow rays, so the interior is lit only through BSDF paths)."""
    sc = cornell_box(light_radiance)
    sc["id"] = "cornell-glass-medium"
    st = np.array([0.8, 1.2, 2.0])
    sc["materials"]["glass"] = {"model": "dielectric", "ior": 1.5,
                                "medium": {"sigma_s": (0.9 * st).tolist(), "sigma_a": (0.1 * st).tolist(), "phase_g": 0.7}}
    sc["objects"].append({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.5, "material": "glass"})
    return sc


def spectral_direct():
    """spectral-direct (tests/witnesses/scenes/spectralFoundation.ts:21-24,33-38): the camera sees
    only a unit grey emitter (D65 white under the spectral lift); with maxBounces 0 every pixel
    projects to (1, 1, 1) through the CIE sensor. The plane is a large quad at z = -1."""
    return {
        "id": "spectral-direct",
        "materials": {"plane": {"model": "lambert", "albedo": 0.0, "emission": 1.0}},
        "objects": [{"type": "quad", "corner": [-50.0, -50.0, -1.0], "edge1": [100.0, 0.0, 0.0],
                     "edge2": [0.0, 100.0, 0.0], "material": "plane"}],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 0.0], "target": [0.0, 0.0, -1.0], "fov": 0.8},
    }


def spectral_furnace(albedo=(0.5, 0.5, 0.5)):
    """spectral-furnace / spectral-colored-furnace (spectralFoundation.ts:15-19,39-48): the
    mesh furnace under the spectral color model; grey albedo 0.5 gives 0.4 (1 - 0.5^7)
    through bounce six, albedo (0.2, 0.5, 0.8) the projected finite series
    (0.22947, 0.39324, 0.76522)."""
    sc = mesh_furnace()
    sc["id"] = "spectral-furnace"
    sc["materials"]["furnace"]["albedo"] = list(albedo)
    return sc


def spectral_slab():
    """spectral-slab (tests/witnesses/scenes/spectralMedia.ts:6-9,41-43): F-SLAB with
    sigma_a = (0.3, 0.7, 1.2) lifted as a coefficient spectrum; the D65 screen seen through
    one unit of ink projects to (0.72520, 0.49546, 0.29589)."""
    sc = slab_witness()
    sc["id"] = "spectral-slab"
    sc["materials"]["ink"] = {"model": "none", "medium": {"sigma_a": [0.3, 0.7, 1.2]}}
    return sc


def cornell_spectral_medium(kelvin=6500.0, scale=12.0):
    """The spectral homogeneous-media workload: the glass-medium Cornell box lit by a
    blackbody quad (kelvin, scale) under the spectral color model (four hero wavelengths)."""
    sc = cornell_glass_medium()
    sc["id"] = "cornell-spectral-medium"
    for l in sc["lights"]:
        l["radiance"] = {"blackbody": {"kelvin": kelvin, "scale": scale}}
    return sc


# ------------------------------------------------ differentiable media inverse render (C4)

INVERSE_N_ALBEDO = 16


def inverse_media_scene(theta):
    """The glass-medium Cornell box with its walls and inner boxes split into 16 grey Lambert
    patches (floor 2 x 2, back wall 2 x 2, left / right / ceiling 2 each, one per inner box);
    theta = [sigma_s RGB, sigma_a RGB, g, albedo x 16] (the adjoint's parameter layout), the
    sphere's interior medium carrying sigma_s / sigma_a / g. RGB light (12, 12, 12)."""
    theta = [float(v) for v in theta]
    assert len(theta) == 7 + INVERSE_N_ALBEDO
    patches = []
    for i in range(2):
        for j in range(2):
            patches.append(_quad_tris((-1 + i, -1, 1 - j), (1, 0, 0), (0, 0, -1)))            # floor
    for i in range(2):
        for j in range(2):
            patches.append(_quad_tris((-1 + i, -1 + j, -1), (1, 0, 0), (0, 1, 0)))            # back
    for j in range(2):
        patches.append(_quad_tris((-1, -1 + j, 1), (0, 0, -2), (0, 1, 0)))                    # left
    for j in range(2):
        patches.append(_quad_tris((1, -1 + j, -1), (0, 0, 2), (0, 1, 0)))                     # right
    for i in range(2):
        patches.append(_quad_tris((-1 + i, 1, -1), (1, 0, 0), (0, 0, 2)))                     # ceiling
    materials, objects = {}, []
    for k, tris in enumerate(patches):
        materials[f"patch{k}"] = {"model": "lambert", "albedo": theta[7 + k]}
        objects.append({"type": "mesh", "triangles": np.asarray(tris), "material": f"patch{k}", "closed": False})
    for k, (lo, hi) in enumerate((((0.05, -1.0, 0.0), (0.65, -0.4, 0.6)), ((-0.65, -1.0, -0.6), (-0.05, 0.2, 0.0)))):
        name = f"patch{14 + k}"
        materials[name] = {"model": "lambert", "albedo": theta[21 + k]}
        objects.append({"type": "mesh", "triangles": _box_tris(lo, hi), "material": name, "closed": True})
    materials["glass"] = {"model": "dielectric", "ior": 1.5,
                          "medium": {"sigma_s": theta[0:3], "sigma_a": theta[3:6], "phase_g": theta[6]}}
    objects.append({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.5, "material": "glass"})
    return {
        "id": "inverse-media",
        "materials": materials,
        "objects": objects,
        "lights": [{"type": "quad", "corner": [-0.25, 0.999, -0.25], "edge1": [0.5, 0.0, 0.0],
                    "edge2": [0.0, 0.0, 0.5], "radiance": [12.0, 12.0, 12.0]}],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 3.5], "target": [0.0, 0.0, 0.0],
                   "fov": math.radians(40.0)},
    }


def inverse_media_params(seed=2024):
    """(hidden target, perturbed init) parameter vectors drawn from a seed: sigma_t per channel
    in [0.6, 2.2] with albedo 0.9 (sigma_s = 0.9 sigma_t), g in [0.3, 0.8], patch albedos in
    [0.2, 0.9]; the init scales every parameter by a factor in [0.7, 1.3] (g and albedos
    clamped to their ranges)."""
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.6, 2.2, size=3)
    target = np.concatenate([0.9 * st, 0.1 * st, [rng.uniform(0.3, 0.8)], rng.uniform(0.2, 0.9, size=INVERSE_N_ALBEDO)])
    init = target * rng.uniform(0.7, 1.3, size=target.size)
    init[6] = float(np.clip(init[6], 0.05, 0.9))
    init[7:] = np.clip(init[7:], 0.05, 0.95)
    return target, init


def inverse_adjoint_materials(comp):
    """(medium material index, albedo material indices) of a compiled inverse_media_scene."""
    names = comp.material_names
    return names.index("glass"), [names.index(f"patch{k}") for k in range(INVERSE_N_ALBEDO)]


# --------------------------------------------------- batched wavefront workload (C3)

def batch_workload(n_scenes=64, seed=99, tri_range=(1000, 500000), res_range=(128, 1024), spp_range=(4, 256)):
    """Seeded scenes of the icosphere-cloud kind with log-uniform triangle count, square
    resolution and spp: returns [(scene, width, height, spp)]. Triangle counts are whole
    detail-4 icospheres (500 triangles each); the cloud's box scales with the cube root of
    the sphere count so density stays comparable."""
    rng = np.random.default_rng(seed)
    out = []
    lg = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    for k in range(n_scenes):
        tris = lg(*tri_range)
        n_spheres = max(2, int(round(tris / 500.0)))
        res = int(round(lg(*res_range)))
        spp = int(round(lg(*spp_range)))
        box = 100.0 * (n_spheres / 2000.0) ** (1.0 / 3.0)
        sc = icosphere_cloud(n_spheres=n_spheres, seed=1000 + k, box=box)
        sc["id"] = f"batch-{k}"
        out.append((sc, res, res, spp))
    return out


def haze():
    """haze (tests/witnesses/scenes/mediaWitness.ts:113-147): a grey room (x, z in [-2.5, 2.5] x
    [-2.5, 6], y in [0, 2.5]) filled with an ambient haze (sigma_a 0.02, sigma_s 0.4, g 0.7), a
    point light (intensity 10) near the ceiling and a small emissive panel on the back wall;
    the camera at (0, 1.2, 5) looks at (0, 1.2, -1). The vertex and equiangular medium-NEE
    placements must converge to the same image."""
    room = _box_tris((-2.5, 0.0, -2.5), (2.5, 2.5, 6.0))[:, [0, 2, 1]]       # inward winding
    return {
        "id": "haze",
        "materials": {"gray": {"model": "lambert", "albedo": 0.5},
                      "panel": {"model": "lambert", "albedo": 0.0, "emission": 4.0},
                      "fog": {"model": "none", "medium": {"sigma_a": 0.02, "sigma_s": 0.4, "phase_g": 0.7}}},
        "objects": [
            {"type": "mesh", "triangles": room, "material": "gray", "closed": False},
            {"type": "mesh", "triangles": _box_tris((-1.6, 0.8, -2.45), (-0.8, 1.6, -2.35)), "material": "panel",
             "closed": True},
        ],
        "lights": [{"type": "point", "position": [0.8, 2.2, -1.0], "intensity": 10.0}],
        "environment": {"type": "none"},
        "ambient_medium": "fog",
        "camera": {"type": "pinhole", "position": [0.0, 1.2, 5.0], "target": [0.0, 1.2, -1.0], "fov": 0.9},
    }


def instanced_cloud(n_instances=2000, seed=7, box=100.0, detail=4, n_materials=3):
    """The icosphere cloud through an instance batch per material kind: one noise-displaced
    detail-4 icosphere prototype (500 triangles) placed n times with seeded rotations,
    lognormal(0, 0.5) scales and positions uniform in [0, box]^3 (GGX, dielectric and Lambert
    batches, 60 / 20 / 20 %); procedural sky."""
    rng = np.random.default_rng(seed)
    V, F = unit_icosphere(detail)
    noise = _value_noise(V * 2.5 + 11.0, seed)
    P = V * (1.0 + 0.15 * (2.0 * noise - 1.0))[:, None]
    tris = P[F]
    kinds = rng.choice(3, size=n_instances, p=[0.6, 0.2, 0.2])
    materials = {"ggx": {"model": "ggx", "f0": [0.9, 0.7, 0.5], "roughness": 0.3},
                 "glass": {"model": "dielectric", "ior": 1.5},
                 "matte": {"model": "lambert", "albedo": 0.6}}
    names = ["ggx", "glass", "matte"]
    placements = [[] for _ in range(3)]
    for k in range(n_instances):
        q = rng.normal(size=4)
        placements[kinds[k]].append({"rotation": (q / np.linalg.norm(q)).tolist(),
                                     "translation": rng.uniform(0.0, box, size=3).tolist(),
                                     "scale": float(rng.lognormal(0.0, 0.5))})
    objects = [{"type": "instances", "triangles": tris, "material": names[m], "closed": True, "placements": placements[m]}
               for m in range(3) if placements[m]]
    return {
        "id": f"instanced-cloud-{n_instances}",
        "materials": materials,
        "objects": objects,
        "lights": [],
        "environment": {"type": "procedural", "intensity": 1.0},
        "camera": {"type": "pinhole", "position": [box * 0.5, box * 0.5, -0.3 * box],
                   "target": [box * 0.5, box * 0.5, box * 0.5], "fov": math.radians(60.0)},
    }


def cornell_identity_instances():
    """The Cornell box with its two inner boxes as instance batches of one identity placement
    each: world-space prototype triangles, rotation (0, 0, 0, 1), translation 0, scale 1. The
    placement transform is then exact, so every hit, normal and containment query equals the
    plain mesh scene's bit for bit."""
    sc = cornell_box()
    for o in sc["objects"]:
        if o["type"] == "mesh" and o["closed"]:
            o["type"] = "instances"
            o["placements"] = [{"rotation": [0.0, 0.0, 0.0, 1.0], "translation": [0.0, 0.0, 0.0], "scale": 1.0}]
    sc["id"] = "cornell-identity-instances"
    return sc


def cornell_glass_instances(n=3, seed=5, detail=2):
    """Cornell box plus a glass instance batch: a detail-2 icosphere placed n times with seeded
    rotations, scales in [0.2, 0.32] and centres inside the box, one of them cutting into the
    tall inner box (nested closed regions and dielectric interfaces through placements)."""
    sc = cornell_box()
    rng = np.random.default_rng(seed)
    V, F = unit_icosphere(detail)
    centres = [[-0.35, 0.25, -0.3]] + [[rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.5), rng.uniform(-0.2, 0.6)]
                                       for _ in range(n - 1)]
    pls = []
    for c in centres:
        q = rng.normal(size=4)
        pls.append({"rotation": (q / np.linalg.norm(q)).tolist(), "translation": list(map(float, c)),
                    "scale": float(rng.uniform(0.2, 0.32))})
    sc["materials"]["glass"] = {"model": "dielectric", "ior": 1.5}
    sc["objects"].append({"type": "instances", "triangles": V[F], "material": "glass", "closed": True, "placements": pls})
    sc["id"] = f"cornell-glass-instances-{n}"
    return sc


def transformed_triangles(tris, rotation, translation, scale):
    """World triangles of a placement in f64: R(q) (s p) + t, q = (x, y, z, w) local -> world."""
    q = np.asarray(rotation, np.float64)
    q = q / np.linalg.norm(q)
    u, w = q[:3], q[3]
    p = np.asarray(tris, np.float64).reshape(-1, 3) * float(scale)
    c = np.cross(u, p) + w * p
    r = p + 2.0 * np.cross(u, c)
    return (r + np.asarray(translation, np.float64)).reshape(-1, 9)


# ------------------------------------------------------------ medium emission witnesses

def emit_slab(sigma_a=1.0, emission=(0.5, 1.0, 2.0)):
    """F-EMIT (tests/witnesses/scenes/emissionWitness.ts:22-56): the F-SLAB geometry with a
    glowing absorbing ink (volume emission coefficient eps) in front of the L = 1 screen:
    L = eps / sigma_a (1 - e^{-sigma_a}) + e^{-sigma_a} = (0.68394, 1, 1.63212) at the centre; the
    green channel is the equilibrium eps = sigma_a L_back, identically 1."""
    sc = slab_witness()
    sc["id"] = "emit"
    sc["materials"]["ink"] = {"model": "none", "medium": {"sigma_a": sigma_a, "emission": list(emission)
                                                          if hasattr(emission, "__len__") else emission}}
    return sc


def emit_sat():
    """F-EMIT-SAT (emissionWitness.ts:78-105): the camera deep inside a uniform glowing scattering
    ambient medium (sigma_a 2, sigma_s 1, eps (1, 2, 4), isotropic): every pixel converges to
    eps / sigma_a = (0.5, 1, 2); at maxBounces N it is (eps / sigma_t) sum_{n<=N} 3^-n."""
    return {
        "id": "emit-sat",
        "materials": {"gray": {"model": "lambert", "albedo": 0.5},
                      "glow": {"model": "none", "medium": {"sigma_a": 2.0, "sigma_s": 1.0, "emission": [1.0, 2.0, 4.0],
                                                           "phase_g": 0.0}}},
        "objects": [{"type": "sphere", "center": [0.0, 1.0, 0.0], "radius": 0.5, "material": "gray"}],
        "lights": [],
        "environment": {"type": "none"},
        "ambient_medium": "glow",
        "camera": {"type": "pinhole", "position": [30.0, 1.0, 0.0], "target": [31.0, 1.0, 0.0], "fov": 0.8},
    }


def emit_scatter(sigma_a=(0.1, 0.1, 0.1), sigma_s=(0.6, 0.6, 0.6)):
    """emit-scatter (emissionWitness.ts:153-188): a glowing scattering fog box (eps (0.8, 0.4, 0.2),
    g 0.3) in a grey room (floor y = 0, ceiling y = 2.5, back wall z = -3 as 100 x 100 quads)
    under a downward ceiling quad light (emission 12): pt-nee and pt estimate the same image
    (volumes are never light-sampled)."""
    return {
        "id": "emit-scatter",
        "materials": {"gray": {"model": "lambert", "albedo": 0.5},
                      "glowfog": {"model": "none", "medium": {"sigma_a": list(sigma_a), "sigma_s": list(sigma_s),
                                                              "emission": [0.8, 0.4, 0.2], "phase_g": 0.3}}},
        "objects": [
            {"type": "quad", "corner": [-50.0, 0.0, -50.0], "edge1": [0.0, 0.0, 100.0], "edge2": [100.0, 0.0, 0.0], "material": "gray"},
            {"type": "quad", "corner": [-50.0, 2.5, -50.0], "edge1": [100.0, 0.0, 0.0], "edge2": [0.0, 0.0, 100.0], "material": "gray"},
            {"type": "quad", "corner": [-50.0, -50.0, -3.0], "edge1": [100.0, 0.0, 0.0], "edge2": [0.0, 100.0, 0.0], "material": "gray"},
            {"type": "mesh", "triangles": _box_tris((-1.0, 0.0, -1.5), (1.0, 2.0, 0.5)), "material": "glowfog", "closed": True},
        ],
        "lights": [{"type": "quad", "corner": [-0.5, 2.48, -1.0], "edge1": [1.0, 0.0, 0.0], "edge2": [0.0, 0.0, 1.0],
                    "radiance": [12.0, 12.0, 12.0]}],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 1.2, 4.2], "target": [0.0, 1.0, 0.0], "fov": 0.9},
    }


# ------------------------------------------------------------------ dispersion witnesses

N_BK7 = {"sellmeier": {"B": [1.03961212, 0.231792344, 1.01046945], "C": [0.00600069867, 0.0200179144, 103.560653]}}


def dispersion_slab(ior=None):
    """spectral-dispersion-slab (tests/witnesses/scenes/spectralDispersion.ts:5-22): a glass slab
    z in [-0.5, 0.5] (Cauchy [1.5, 0.01] by default) in front of an emission-1 screen at z = -2,
    seen straight on: two normal transmissions, (1 - F(lambda))^2 of normalized D65 ->
    (0.91574, 0.91323, 0.90889). A second hero collapse would quadruple the answer."""
    return {
        "id": "spectral-dispersion-slab",
        "materials": {"glass": {"model": "dielectric", "ior": ior if ior is not None else {"cauchy": [1.5, 0.01]},
                                "transmittance": 1.0},
                      "lamp": {"model": "lambert", "albedo": 0.0, "emission": 1.0}},
        "objects": [
            {"type": "mesh", "triangles": _box_tris((-100.0, -100.0, -0.5), (100.0, 100.0, 0.5)), "material": "glass", "closed": True},
            {"type": "quad", "corner": [-200.0, -200.0, -2.0], "edge1": [400.0, 0.0, 0.0], "edge2": [0.0, 400.0, 0.0], "material": "lamp"},
        ],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 2.0], "target": [0.0, 0.0, -2.0], "fov": 0.3},
    }


def dispersion_prism(ior=None):
    """spectral-dispersion-prism (spectralDispersion.ts:24-42): a tall closed triangular prism
    (Cauchy [1.45, 0.02] by default) before a thin emissive stripe (emission 3) and a dark
    surround wall; the stripe's image through the prism splits by wavelength."""
    P = np.array([[-1, -1, -0.7], [1, -1, -0.7], [0, -1, 0.9], [-1, 1, -0.7], [1, 1, -0.7], [0, 1, 0.9]], np.float64)
    I = np.array([[0, 1, 2], [3, 5, 4], [0, 3, 4], [0, 4, 1], [1, 4, 5], [1, 5, 2], [2, 5, 3], [2, 3, 0]])
    return {
        "id": "spectral-dispersion-prism",
        "materials": {"glass": {"model": "dielectric", "ior": ior if ior is not None else {"cauchy": [1.45, 0.02]}},
                      "lamp": {"model": "lambert", "albedo": 0.0, "emission": 3.0},
                      "surround": {"model": "lambert", "albedo": 0.1}},
        "objects": [
            {"type": "mesh", "triangles": P[I], "material": "glass", "closed": True},
            {"type": "mesh", "triangles": _box_tris((-0.15, -3.0, -3.03), (0.15, 3.0, -2.97)), "material": "lamp", "closed": True},
            {"type": "quad", "corner": [-50.0, -50.0, -3.1], "edge1": [100.0, 0.0, 0.0], "edge2": [0.0, 100.0, 0.0], "material": "surround"},
        ],
        "lights": [],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.3, 5.0], "target": [0.0, 0.0, 0.0], "fov": math.radians(38.0)},
    }


def smooth_sphere(mode="smooth", detail=3):
    """A Lambert unit sphere lit by one point light, seen from the front: mode 'smooth' is the
    icosphere mesh with per-corner normals = the unit vertex directions (MeshObject.normals,
    mesh.glsl:84-109), 'flat' the same mesh with geometric normals, 'analytic' the exact sphere.
    Smooth shading must approach the analytic image far closer than flat facets do."""
    V, F = unit_icosphere(detail)
    obj = ({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 1.0, "material": "m"} if mode == "analytic" else
           {"type": "mesh", "triangles": V[F], "material": "m", "closed": True})
    if mode == "smooth":
        obj["normals"] = (V / np.linalg.norm(V, axis=1, keepdims=True))[F]
    return {
        "id": f"sphere-{mode}",
        "materials": {"m": {"model": "lambert", "albedo": 0.8}},
        "objects": [obj],
        "lights": [{"type": "point", "position": [2.0, 3.0, 4.0], "intensity": 30.0}],
        "environment": {"type": "none"},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 4.0], "target": [0.0, 0.0, 0.0], "fov": 0.6},
    }


def cornell_smooth_instances(n=4, seed=11, detail=2):
    """Cornell box with a smooth-shaded GGX instance batch (icosphere prototype with per-corner
    normals) at seeded rotations, scales and positions, plus one smooth glass icosphere mesh."""
    sc = cornell_box()
    rng = np.random.default_rng(seed)
    V, F = unit_icosphere(detail)
    N = (V / np.linalg.norm(V, axis=1, keepdims=True))[F]
    pls = []
    for _ in range(n):
        q = rng.normal(size=4)
        pls.append({"rotation": (q / np.linalg.norm(q)).tolist(),
                    "translation": [float(rng.uniform(-0.6, 0.6)), float(rng.uniform(-0.3, 0.6)), float(rng.uniform(-0.5, 0.5))],
                    "scale": float(rng.uniform(0.12, 0.25))})
    sc["materials"]["metal"] = {"model": "ggx", "f0": [0.9, 0.8, 0.6], "roughness": 0.25}
    sc["materials"]["glass"] = {"model": "dielectric", "ior": 1.5}
    sc["objects"].append({"type": "instances", "triangles": V[F], "normals": N, "material": "metal", "closed": True,
                          "placements": pls})
    sc["objects"].append({"type": "mesh", "triangles": V[F] * 0.3 + np.array([0.35, -0.7, 0.4]), "normals": N,
                          "material": "glass", "closed": True})
    sc["id"] = "cornell-smooth-instances"
    return sc


def furnace_sky(sample_as_light=True, color=(1.0, 1.0, 1.0)):
    """furnace-sky (tests/witnesses/scenes/envScenes.ts:69-120): a Lambert sphere (albedo 0.4)
    under a uniform constant sky L = 1, sampled as a light. A convex body never re-illuminates
    itself, so the sphere's pixels are rho L = 0.4 and the sky's pixels are L = 1 exactly, for
    every estimator."""
    return {
        "id": "furnace-sky",
        "materials": {"gray": {"model": "lambert", "albedo": [0.4, 0.4, 0.4]}},
        "objects": [{"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 1.0, "material": "gray"}],
        "lights": [],
        "environment": {"type": "constant", "color": list(color), "intensity": 1.0, "sample_as_light": sample_as_light},
        "camera": {"type": "pinhole", "position": [0.0, 0.0, 3.5], "target": [0.0, 0.0, 0.0], "fov": 0.8},
    }


def rough_glass_furnace(roughness=0.03):
    """furnace-sky with a rough dielectric sphere (index 1.5, clear): every camera path enters
    and leaves the sphere, so eta^2 cancels; near-smooth (alpha at the 1e-3 floor) the sphere
    renders L = 1 up to the microfacet model's single-scattering loss."""
    sc = furnace_sky(True)
    sc["id"] = "rough-glass-furnace"
    sc["materials"] = {"frost": {"model": "rough_dielectric", "ior": 1.5, "roughness": roughness}}
    sc["objects"][0]["material"] = "frost"
    return sc


def cornell_rough_glass(roughness=0.3, medium=None):
    """The Cornell box with a rough glass sphere (radius 0.3) and the short inner box as a rough
    glass mesh, optionally filled with a homogeneous medium."""
    sc = cornell_box()
    sc["id"] = "cornell-rough-glass"
    mat = {"model": "rough_dielectric", "ior": 1.5, "roughness": roughness, "transmittance": [1.0, 0.95, 0.9]}
    if medium is not None:
        mat["medium"] = medium
    sc["materials"]["frost"] = mat
    sc["objects"][3]["material"] = "frost"
    sc["objects"].append({"type": "sphere", "center": [-0.45, 0.35, 0.35], "radius": 0.3, "material": "frost"})
    return sc


def mirror_furnace(f0=(1.0, 1.0, 1.0)):
    """furnace-sky with the reference's mirror material on the sphere: one delta reflection per
    camera ray (a convex mirror never sees itself), weight Schlick(f0, cos), then the L = 1 sky;
    f0 = 1 renders the sphere as L = 1 exactly."""
    sc = furnace_sky(True)
    sc["id"] = "mirror-furnace"
    sc["materials"] = {"chrome": {"model": "mirror", "f0": list(f0)}}
    sc["objects"][0]["material"] = "chrome"
    return sc


def cornell_mirror(f0=(0.9, 0.9, 0.9)):
    """The Cornell box with the tall inner box turned into a mirror mesh and a mirror sphere
    (radius 0.3) on the floor: delta reflections that see the light and each other."""
    sc = cornell_box()
    sc["id"] = "cornell-mirror"
    sc["materials"]["chrome"] = {"model": "mirror", "f0": list(f0)}
    sc["objects"][4]["material"] = "chrome"
    sc["objects"].append({"type": "sphere", "center": [0.45, -0.7, 0.35], "radius": 0.3, "material": "chrome"})
    return sc


def thinlens(sc, aperture=0.05, focus_distance=3.0):
    """The scene seen through the reference's thin-lens camera (camera/thinlens/thinlens.ts:
    aperture = lens radius, focusDistance along forward), with the scene camera's pose and fov."""
    sc = dict(sc)
    cam = dict(sc["camera"])
    cam.update({"type": "thinlens", "aperture": aperture, "focusDistance": focus_distance})
    sc["camera"] = cam
    sc["id"] = sc["id"] + "-thinlens"
    return sc


def orthographic(sc, scale=0.5):
    """The scene seen through the reference's orthographic camera (u_orthoScale = scale), as the
    media witnesses pose it (tests/witnesses/scenes/spectralMedia.ts:10-16)."""
    sc = dict(sc)
    cam = dict(sc["camera"])
    cam.pop("fov", None)
    cam.update({"type": "orthographic", "scale": scale})
    sc["camera"] = cam
    sc["id"] = sc["id"] + "-ortho"
    return sc
