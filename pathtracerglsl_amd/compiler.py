"""Host compiler: SceneDescription + RenderStrategy -> CompiledRenderer for the MI355X backend.

Mirrors the reference pipeline Analyze -> Validate -> Plan (src/compiler/Compiler.ts:14-72,
src/compiler/analyze/Analyzer.ts:10, src/compiler/analyze/Validator.ts:50,
src/compiler/plan/Planner.ts:72). The Generate stage is replaced: instead of emitting GLSL, the
plan is lowered to a pt_program record (template selection on the device side) plus the scene
arrays of the data rail and the host-derived constants the generator would bake as literals
(camera basis, tan(fov/2), quad normals/areas, light selection pdfs — all f64 on the host,
rounded once to f32, as the reference's formatFloat literals are).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi, spectral

ESTIMATORS = {"none": abi.EST_PT, "pt": abi.EST_PT, "nee": abi.EST_NEE, "mis": abi.EST_MIS}
MODELS = {"lambert": abi.MAT_LAMBERT, "dielectric": abi.MAT_DIELECTRIC, "ggx": abi.MAT_GGX, "none": abi.MAT_NULL,
          "mirror": abi.MAT_MIRROR, "rough_dielectric": abi.MAT_ROUGH_DIELECTRIC, "checker": abi.MAT_CHECKER}
# models whose emission term is identically zero (their descriptors' emissive capability is false)
NON_EMISSIVE = ("mirror", "rough_dielectric", "checker")
# accumulators (components/accumulator/index.ts registry): estimator.accumulation.type
ACCUMULATIONS = {"average": abi.ACC_AVERAGE, "oneshot": abi.ACC_ONESHOT}
# volume scattering models (components/volume_scattering/index.ts registry): medium.model
PHASE_MODELS = {"hg": abi.PHASE_HG, "rayleigh": abi.PHASE_RAYLEIGH}


class CompileError(ValueError):
    """Validation failure (the reference's DiagnosticBag errors, src/errors/core/DiagnosticBag.ts)."""


@dataclass
class RenderStrategy:
    """measurement/estimator knobs (src/compiler/types.ts:561-721)."""
    direct_lighting: str = "mis"          # 'none' | 'nee' | 'mis'
    max_bounces: int = 8
    rr_start: int = 3                     # russianRoulette.startDepth (-1 = off)
    rr_max_survival: float = 0.95
    frames_in_flight: int = 1
    max_null_crossings: int = 32          # measurement.maxNullCrossings (plan/measurement.ts:12)
    color: str = "rgb"                    # measurement.color: 'rgb' | 'spectral'
    wavelengths: str = "visible"          # estimator.wavelengths (spectral): 'visible' | 'uniform'
    medium_light_sampling: str = "vertex" # estimator.mediumLightSampling: 'vertex' | 'equiangular'
    variance: bool = False                # accumulate the variance statistic picture
    aovs: tuple = ()                      # measurement.aovs: first-hit guides 'albedo' / 'normal' (RGB)
    volume_sampling: str = "analytic"     # estimator.volumeSampling: 'analytic' | 'delta-tracking'
    scattering: str = "full"              # measurement.scattering: 'full' | 'ignored' (absorbing-only media)
    light_selection: str = "power"        # estimator.lightSelection: 'power' | 'uniform' | 'bvh'
    accumulation: str = "average"         # estimator.accumulation.type: 'average' | 'oneshot'


@dataclass
class CompiledRenderer:
    id: str
    scene: "abi.SceneArrays"
    program: "abi.PtProgram"
    width: int
    height: int
    camera: dict
    reset_salt: int = 1234
    warnings: list = field(default_factory=list)
    material_names: list = field(default_factory=list)   # authored material of each record index
    source: dict = None                  # the authored scene (live parameters recompile from it)
    strategy: "RenderStrategy" = None

    def params(self, reset_salt=None):
        return camera_params(self.camera, self.width, self.height,
                             self.reset_salt if reset_salt is None else reset_salt)


def camera_basis(position, target):
    """cameraBasis (src/components/camera/basis.ts:50-56), in f64."""
    p = np.asarray(position, np.float64)
    t = np.asarray(target, np.float64)
    f = t - p
    f = f / math.hypot(*f)
    up_ref = np.array([0.0, 0.0, 1.0]) if abs(f[1]) > 0.999999 else np.array([0.0, 1.0, 0.0])
    r = np.cross(f, up_ref)
    r = r / math.hypot(*r)
    u = np.cross(r, f)
    return f, r, u


def camera_params(camera, width, height, reset_salt=1234):
    f, r, u = camera_basis(camera["position"], camera["target"])
    prm = abi.PtParams()
    for dst, src in ((prm.cam_position, camera["position"]), (prm.cam_forward, f),
                     (prm.cam_right, r), (prm.cam_up, u)):
        for i in range(3):
            dst[i] = float(np.float32(src[i]))
    if camera.get("type", "pinhole") == "orthographic":
        prm.camera = abi.CAM_ORTHOGRAPHIC
        prm.ortho_scale = float(np.float32(camera.get("scale", 1.0)))     # u_orthoScale
    elif camera.get("type") == "fisheye":                                # fisheye.ts fisheyeK, in f64
        prm.camera = abi.CAM_FISHEYE
        proj = camera["projection"]
        tmax = float(camera["fov"]) * 0.5
        k = {"equisolid": math.sin(tmax * 0.5), "stereographic": math.tan(tmax * 0.5),
             "orthographic": math.sin(tmax)}.get(proj, tmax)
        prm.projection = abi.FISHEYE_PROJECTIONS[proj]
        prm.fisheye_k = float(np.float32(k))
    elif camera.get("type") == "cylindrical":                            # cylindrical.ts u_cylFocal
        prm.camera = abi.CAM_CYLINDRICAL
        prm.cyl_focal = float(np.float32(width / max(float(camera["hfov"]) * math.pi / 180.0, 1e-6)))
    elif camera.get("type") == "equirect":
        prm.camera = abi.CAM_EQUIRECT
    else:
        prm.camera = abi.CAM_PINHOLE
        prm.tan_fov = float(np.float32(math.tan(camera["fov"] / 2.0)))   # features/camera.ts:54-66
        if camera.get("type") == "thinlens":                             # u_aperture, u_focusDistance
            prm.camera = abi.CAM_THINLENS
            prm.aperture = float(np.float32(camera["aperture"]))
            prm.focus_distance = float(np.float32(camera.get("focusDistance", camera.get("focus_distance"))))
    prm.aspect = float(np.float32(width / height))
    prm.reset_salt = reset_salt
    return prm


INDEX_REFERENCE_NM = 587.5618          # components/color/spectra/spectrumKind.ts:15
SPECTRAL_RANGE_NM = (360.0, 830.0)


def cauchy_index(coef, lam_nm):
    """cauchyIndex (components/color/spectra/cauchy/cauchy.ts): A + B / l^2 + C / l^4, l in um."""
    a, b, c = (list(coef) + [0.0])[:3]
    inv2 = (1000.0 / lam_nm) ** 2
    return a + inv2 * (b + inv2 * c)


def sellmeier_index(B, C, lam_nm):
    """sellmeierIndex (components/color/spectra/sellmeier/sellmeier.ts): three-term, l in um."""
    l2 = (lam_nm / 1000.0) ** 2
    return math.sqrt(1.0 + sum(b * l2 / (l2 - c) for b, c in zip(B, C)))


def index_record(value, spectral_program, label):
    """(ior at 587.5618 nm, PT_IOR_* kind, 8 coefficients) of an index property: a number, or
    {"cauchy": [A, B, C?]} / {"sellmeier": {"B": [..], "C": [..]}} with the kinds' validation.
    The dispersive formula is kept only for spectral programs (RGB evaluates the reference index);
    a formula constant in wavelength folds to its value, as the kinds' emit() does."""
    coef = [0.0] * 8
    if isinstance(value, dict) and "cauchy" in value:
        c = value["cauchy"]
        if (not isinstance(c, (list, tuple)) or not 2 <= len(c) <= 3 or any(not math.isfinite(float(x)) for x in c)
                or not float(c[0]) > 0 or any(float(x) < 0 for x in c[1:])):
            raise CompileError(f"{label}: Cauchy requires constant [A, B, C?], with finite A > 0 and B, C >= 0 (micrometre units)")
        c = [float(x) for x in c] + [0.0] * (3 - len(c))
        n = cauchy_index(c, INDEX_REFERENCE_NM)
        dispersive = c[1] != 0.0 or c[2] != 0.0
        coef[:3] = c
        return n, (abi.IOR_CAUCHY if spectral_program and dispersive else abi.IOR_CONSTANT), coef
    if isinstance(value, dict) and "sellmeier" in value:
        B, C = [float(x) for x in value["sellmeier"]["B"]], [float(x) for x in value["sellmeier"]["C"]]
        if len(B) != 3 or len(C) != 3 or any(x < 0 or not math.isfinite(x) for x in B + C):
            raise CompileError(f"{label}: Sellmeier requires three non-negative B and C terms")
        lo2, hi2 = (SPECTRAL_RANGE_NM[0] / 1000.0) ** 2, (SPECTRAL_RANGE_NM[1] / 1000.0) ** 2
        if any(lo2 <= c <= hi2 for c in C):
            raise CompileError(f"{label}: Sellmeier poles must lie outside 360–830 nm")
        try:
            ok = sellmeier_index(B, C, SPECTRAL_RANGE_NM[1]) > 0 and math.isfinite(sellmeier_index(B, C, SPECTRAL_RANGE_NM[0]))
        except ValueError:
            ok = False
        if not ok:
            raise CompileError(f"{label}: Sellmeier must give finite positive n² throughout the spectral interval")
        n = sellmeier_index(B, C, INDEX_REFERENCE_NM)
        dispersive = any(b != 0.0 and c != 0.0 for b, c in zip(B, C))
        coef[:6] = B + C
        return n, (abi.IOR_SELLMEIER if spectral_program and dispersive else abi.IOR_CONSTANT), coef
    return float(value), abi.IOR_CONSTANT, coef


def scene_radius_estimate(scene):
    """sceneRadiusEstimate (features/lighting.ts:459-474): max(1, |point| + sum of lengths) over
    the primitive objects (sphere: centre + radius; quad: corner) and the lights (quad corner,
    point position); meshes are not primitive objects and do not count."""
    r = 1.0
    for o in scene.get("objects", []):
        if o["type"] == "sphere":
            r = max(r, math.hypot(*o["center"]) + float(o["radius"]))
        elif o["type"] == "quad":
            r = max(r, math.hypot(*o["corner"]))
    for l in scene.get("lights", []):
        r = max(r, math.hypot(*(l["corner"] if l["type"] == "quad" else l["position"])))
    return r


ENV_CHARTS = {"equirect": abi.CHART_EQUIRECT, "octahedral": abi.CHART_OCTAHEDRAL}


def select_pdfs(light_powers, selection):
    """computeSelectPdf + cumulative (features/lighting.ts:375-380, 537-542): weight 1 per light
    under 'uniform', its power under 'power'; pdf_i = w_i / sum(w) and cdf_i the running sum
    of those pdfs, in f64 (the reference bakes both as literals, so the device sees them rounded
    to fp32)."""
    weights = [1.0 if selection == "uniform" else float(p) for p in light_powers]
    total = 0.0
    for w in weights:
        total += w
    pdfs = [w / total for w in weights]
    cdf, acc = [], 0.0
    for p in pdfs:
        acc += p
        cdf.append(acc)
    return pdfs, cdf


def env_select_probability(scene, env, light_powers, selection="power", total_weight=None):
    """envSelectionProbability (features/lighting.ts:488-518). 'uniform': 1 / (N + 1). 'power':
    Phi_env / (Phi_env + sum Phi_light), Phi_env = 4 pi^2 r^2 mean(L) intensity. Both clamped to
    [0.01, 0.99]. mean(L): the constant colour's channel mean, or the importance table's mass
    totalWeight pi / (2 W H) for the tabulated sky on the equirect chart, totalWeight / N^2 on the
    N x N octahedral chart (lighting.ts:502-512). totalWeight is the env.totalWeight parameter:
    the mass of the table the library bakes, unless a registered table's is given."""
    if selection == "uniform":
        return min(0.99, max(0.01, 1.0 / (len(light_powers) + 1)))
    if selection == "bvh":          # the environment against the whole tree (lighting.ts:477-493)
        return 0.5
    if env.type == abi.ENV_CONSTANT:
        mean_l = (env.color[0] + env.color[1] + env.color[2]) / 3.0
    else:
        from . import _lib
        total = float(total_weight) if total_weight is not None else _lib.load().pt_env_total_weight(env)
        if env.chart == abi.CHART_OCTAHEDRAL:
            mean_l = total / float(env.table_h * env.table_h) if total > 0 else 1.0
        else:
            mean_l = total * math.pi / (2.0 * env.table_w * env.table_h) if total > 0 else 1.0
    r = scene_radius_estimate(scene)
    phi_env = 4.0 * math.pi * math.pi * r * r * mean_l * env.intensity
    phi_fin = sum(light_powers)
    return min(0.99, max(0.01, phi_env / max(phi_env + phi_fin, 1e-20)))


def _rgb3(v):
    return [float(v)] * 3 if not hasattr(v, "__len__") else [float(x) for x in v]


def medium_arm(med, strategy):
    """The arm a constant medium is sampled by (src/compiler/types.ts:351-417: mediumScatters,
    isEmissiveMedium, mediumRoutesToTracking; features/materials.ts:675-728 dispatch)."""
    scatters = strategy.scattering == "full" and any(x != 0.0 for x in _rgb3(med.get("sigma_s", 0.0)))
    emissive = any(x > 0.0 for x in spectral.rgb_of(med.get("emission", 0.0)))
    if scatters and emissive:
        return abi.MED_DELTA
    return abi.MED_ANALYTIC if scatters else abi.MED_ABSORB


def _analyze_validate(scene, strategy):
    if strategy.direct_lighting not in ESTIMATORS:
        raise CompileError(f"unknown directLighting '{strategy.direct_lighting}'")
    if strategy.max_bounces < 0:
        raise CompileError("maxBounces must be >= 0")
    if not (0.0 < strategy.rr_max_survival <= 1.0):
        raise CompileError("russianRoulette.maxSurvival must be in (0, 1]")
    if strategy.color not in ("rgb", "spectral"):
        raise CompileError(f"unknown color model '{strategy.color}'")
    if strategy.accumulation not in ACCUMULATIONS:
        raise CompileError(f"Accumulation type '{strategy.accumulation}' not yet supported")
    if strategy.accumulation == "oneshot" and strategy.variance:
        raise CompileError("Accumulation 'oneshot' does not support variance (it keeps no statistic)")
    if strategy.wavelengths not in ("visible", "uniform"):
        raise CompileError(f"unknown wavelength sampler '{strategy.wavelengths}'")
    if strategy.medium_light_sampling not in ("vertex", "equiangular"):
        raise CompileError(f"unknown mediumLightSampling '{strategy.medium_light_sampling}'")
    if strategy.medium_light_sampling == "equiangular":
        # Validator.ts:506-579: equiangular placement runs under pt-nee with delta lights only
        if strategy.direct_lighting != "nee":
            raise CompileError("mediumLightSampling 'equiangular' requires directLighting 'nee'")
        if not scene.get("lights") or any(l["type"] != "point" for l in scene.get("lights", [])):
            raise CompileError("mediumLightSampling 'equiangular' requires every light to be a point light")
        if scene.get("environment", {"type": "none"}).get("type", "none") != "none":
            raise CompileError("mediumLightSampling 'equiangular' does not combine with an environment light")
    # volumeSampling axis (Validator.ts:497-615): a constant-emission SCATTERING medium routes to
    # the null-collision (delta-tracking) arm; everything else stays on the analytic arms
    if strategy.light_selection not in ("power", "uniform", "bvh"):
        raise CompileError(f"estimator.lightSelection '{strategy.light_selection}' not supported — use 'power', "
                           "'uniform' or 'bvh'")
    if strategy.light_selection == "bvh" and strategy.medium_light_sampling == "equiangular":
        raise CompileError("mediumLightSampling 'equiangular' with lightSelection 'bvh' is not supported by this backend")
    if strategy.scattering not in ("full", "ignored"):
        raise CompileError(f"unknown measurement.scattering '{strategy.scattering}'")
    if strategy.volume_sampling not in ("analytic", "delta-tracking"):
        raise CompileError(f"volumeSampling '{strategy.volume_sampling}' not supported — use 'analytic' (constant "
                           "media, exact) or 'delta-tracking' (null-collision)")
    tracked = [n for n, m in scene.get("materials", {}).items()
               if m.get("medium") is not None and medium_arm(m["medium"], strategy) == abi.MED_DELTA]
    if strategy.volume_sampling == "delta-tracking":
        if strategy.direct_lighting == "mis":
            raise CompileError("volumeSampling 'delta-tracking' with directLighting 'mis' is reserved — delta "
                               "tracking's distance pdf is unknowable; use 'nee' or 'none'")
    elif tracked:
        raise CompileError("This scene has media that need the null-collision arms (an emissive scattering medium: "
                           f"{', '.join(tracked)}) but volumeSampling is 'analytic' — set volumeSampling 'delta-tracking'")
    if tracked and strategy.medium_light_sampling == "equiangular":
        raise CompileError("mediumLightSampling 'equiangular' does not support media on the null-collision arms; use 'vertex'")
    cam = scene.get("camera", {})
    for a in strategy.aovs:
        if a not in ("albedo", "normal"):
            raise CompileError(f"unknown aov '{a}' (known: albedo, normal)")
    if strategy.aovs and strategy.color != "rgb":
        raise CompileError("first-hit guides (aovs) are provided for RGB programs only by this backend")
    if cam.get("type", "pinhole") not in ("pinhole", "orthographic", "thinlens", "fisheye", "cylindrical", "equirect"):
        raise CompileError(f"camera '{scene['camera'].get('type')}' is not supported by this backend")
    if cam.get("type") == "thinlens":     # camera/thinlens/thinlens.ts:15-16
        if not float(cam.get("aperture", -1.0)) >= 0.0:
            raise CompileError("camera 'thinlens': aperture must be nonnegative")
        if not float(cam.get("focusDistance", cam.get("focus_distance", 0.0))) > 0.0:
            raise CompileError("camera 'thinlens': focusDistance must be positive")
    if cam.get("type") == "fisheye":      # camera/fisheye/fisheye.ts:31-53
        proj = cam.get("projection")
        if proj not in abi.FISHEYE_PROJECTIONS:
            raise CompileError(f"camera 'fisheye': projection must be one of {sorted(abi.FISHEYE_PROJECTIONS)}")
        fov = float(cam.get("fov", 0.0))
        top, inclusive = (math.pi, True) if proj == "orthographic" else (2 * math.pi, proj != "stereographic")
        if not (fov > 0.0 and (fov <= top if inclusive else fov < top)):
            raise CompileError(f"fisheye '{proj}' fov must be {'<=' if inclusive else '<'} "
                               f"{'pi' if top == math.pi else '2 pi'} radians (got {fov})")
    if cam.get("type") == "cylindrical" and not float(cam.get("hfov", 0.0)) > 0.0:
        raise CompileError("camera 'cylindrical': hfov (degrees) must be positive")
    mats = scene.get("materials", {})
    for name, m in mats.items():
        if m.get("model") not in MODELS:
            raise CompileError(f"material '{name}': model '{m.get('model')}' is not supported by this backend")
        if spectral.is_gaussian(m.get("emission")):
            raise CompileError(f"material '{name}': emission: gaussian values are reflectances only")
        for k in ("albedo", "f0", "transmittance"):
            if spectral.is_gaussian(m.get(k)):
                try:
                    spectral.gaussian_params(m[k])
                except spectral.SpectralDataError as e:
                    raise CompileError(f"material '{name}': {k}: {e}") from None
        med = m.get("medium")
        if med is not None:
            if med.get("model", "hg") not in PHASE_MODELS:
                raise CompileError(f"material '{name}': medium model '{med.get('model')}' is not a volume scattering "
                                   f"model (known: {', '.join(PHASE_MODELS)})")
            for k in ("sigma_a", "sigma_s", "emission"):
                if spectral.is_gaussian(med.get(k)):
                    raise CompileError(f"material '{name}': medium.{k}: gaussian values are reflectances only")
    for o in scene.get("objects", []):
        if o.get("material") not in mats:
            raise CompileError(f"object references unknown material '{o.get('material')}'")
        if o["type"] not in ("mesh", "sphere", "quad", "instances"):
            raise CompileError(f"object type '{o['type']}' is not supported by this backend")
    env = scene.get("environment", {"type": "none"})
    if env.get("type") not in ("none", "procedural", "constant", "image"):
        raise CompileError(f"environment '{env.get('type')}' is not supported by this backend")
    if env.get("type") == "image":
        img = env.get("image")
        if not (isinstance(img, np.ndarray) and img.ndim == 3 and img.shape[2] == 3 and img.shape[0] >= 1
                and 1 <= img.shape[1] <= 4096):
            raise CompileError("environment 'image': 'image' must be an H x W x 3 array of linear RGB (1 <= W <= 4096)")
        if not (np.all(np.isfinite(img)) and np.all(img >= 0)):
            raise CompileError("Environment RGB samples must be finite and nonnegative")
    if env.get("type") == "procedural":
        sky_body(env)
    if spectral.is_gaussian(env.get("color")):
        raise CompileError("environment.color: gaussian values are reflectances only")
    for l in scene.get("lights", []):
        if spectral.is_gaussian(l.get("radiance", l.get("intensity"))):
            raise CompileError("light emission: gaussian values are reflectances only")
        if l["type"] == "quad":
            area = np.linalg.norm(np.cross(l["edge1"], l["edge2"]))
            if area < 1e-8:
                raise CompileError("quad edges are parallel or near-parallel — area must be >= 1e-8")
        elif l["type"] != "point":
            raise CompileError(f"light '{l['type']}' is not supported by this backend")


def placement_arrays(placements):
    """(q, t, s) as f64 arrays [N, 4], [N, 3], [N] of an instance batch's placements: a list of
    Transform-like records {rotation: unit quaternion x y z w or {axis, angle}, translation or
    position, scale}, or the packed form {count, positions, orientations?, sizes?} an instance
    cloud carries (compiler/types.ts PackedPlacements; identity rotation and scale 1 where a
    column is absent)."""
    if isinstance(placements, dict):
        n = int(placements["count"])
        t = np.asarray(placements["positions"], np.float64).reshape(n, 3)
        q = (np.asarray(placements["orientations"], np.float64).reshape(n, 4) if placements.get("orientations") is not None
             else np.tile([0.0, 0.0, 0.0, 1.0], (n, 1)))
        s = (np.asarray(placements["sizes"], np.float64).reshape(n) if placements.get("sizes") is not None
             else np.ones(n))
        return q, t, s
    n = len(placements)
    q, t, s = np.zeros((n, 4)), np.zeros((n, 3)), np.ones(n)
    for i, pl in enumerate(placements):
        q[i] = rotation_quaternion(pl.get("rotation"))
        t[i] = pl.get("translation", pl.get("position", [0.0, 0.0, 0.0]))
        s[i] = float(pl.get("scale", 1.0))
    return q, t, s


def rotation_quaternion(r):
    """A Transform rotation as a quaternion x y z w (components/geometry/similarity.ts:59-64,
    218-228): a quaternion as given (normalised by the placement lowering), or axis-angle
    {axis, angle} as (axis sin(angle/2) / |axis|, cos(angle/2)); identity when absent."""
    if r is None:
        return [0.0, 0.0, 0.0, 1.0]
    if isinstance(r, dict):
        ax = [float(a) for a in r["axis"]]
        ln = math.hypot(*ax)
        h = float(r["angle"]) / 2.0
        sn = math.sin(h) / ln
        return [ax[0] * sn, ax[1] * sn, ax[2] * sn, math.cos(h)]
    return [float(c) for c in r]


def indexed_mesh_object(o):
    """A MeshObject in the reference's indexed form (compiler/types.ts:134-158: object-local
    positions, indices, optional per-vertex normals, optional transform) as this compiler's
    triangle form. Without a transform it is a world-space mesh; with one it is walked as one
    placement of a local-space prototype, as the reference walks a transformed mesh
    (app/sceneData.ts:72-80). Other objects pass through."""
    if o.get("type") != "mesh" or "positions" not in o:
        return o
    P = np.asarray(o["positions"], np.float32).reshape(-1, 3)
    I = np.asarray(o["indices"], np.int64).reshape(-1, 3)
    if I.size and (I.min() < 0 or I.max() >= P.shape[0]):
        raise CompileError("mesh indices must name vertices of the mesh")
    out = {"type": "mesh", "triangles": P[I].reshape(-1, 9).astype(np.float64), "material": o["material"],
           "closed": bool(o.get("closed", False))}
    if o.get("normals") is not None:
        N = np.asarray(o["normals"], np.float32).reshape(-1, 3)
        out["normals"] = N[I].reshape(-1, 9).astype(np.float64)
    tr = o.get("transform")
    if tr is not None:
        out["type"] = "instances"
        out["placements"] = [{"rotation": rotation_quaternion(tr.get("rotation")),
                              "translation": list(tr.get("position", [0.0, 0.0, 0.0])),
                              "scale": float(tr.get("scale", 1.0))}]
    return out


def compile_scene(scene, strategy=None, width=256, height=256, reset_salt=1234):
    """Compile one (scene, strategy) pair into a CompiledRenderer."""
    strategy = strategy or RenderStrategy()
    _analyze_validate(scene, strategy)

    spec = strategy.color == "spectral"

    def srec(value, role):
        if not spec:
            return np.zeros(1, abi.SPECTRUM_DTYPE)[0]
        try:
            return spectral.spectrum_record(value, role)
        except spectral.SpectralDataError as e:
            raise CompileError(str(e)) from None

    def emits(value):
        if isinstance(value, dict) and "blackbody" in value:
            return float(value["blackbody"].get("scale", 1.0)) > 0.0
        return any(e > 0 for e in spectral.rgb_of(value))

    mat_names = list(scene.get("materials", {}).keys())
    mat_index = {n: i for i, n in enumerate(mat_names)}
    materials = []
    for n in mat_names:
        m = scene["materials"][n]
        rec = np.zeros(1, abi.MATERIAL_DTYPE)[0]
        rec["model"] = MODELS[m["model"]]
        albedo = m.get("albedo", [0.8, 0.8, 0.8]) if m["model"] != "checker" else m.get("albedo_a", 0.9)
        rec["albedo"][:3] = spectral.rgb_of(albedo)
        rec["s_albedo"] = srec(albedo, spectral.ROLE_REFLECTANCE)
        em = m.get("emission", [0.0, 0.0, 0.0]) if m["model"] not in NON_EMISSIVE else [0.0, 0.0, 0.0]
        rec["emission"][:3] = spectral.rgb_of(em)
        rec["s_emission"] = srec(em, spectral.ROLE_EMISSION)
        rec["emissive"] = int(emits(em))
        f0 = m.get("f0", [0.9, 0.9, 0.9]) if m["model"] != "checker" else m.get("albedo_b", 0.1)
        rec["f0"][:3] = spectral.rgb_of(f0)
        if m["model"] == "checker":           # checker.ts: albedo_a 0.9, albedo_b 0.1, uv_scale 8 (positive)
            scale = float(m.get("uv_scale", 8.0))
            if not (scale > 0 and math.isfinite(scale)):
                raise CompileError(f"material '{n}': uv_scale must be positive")
            rec["pad1"] = scale
        rec["s_f0"] = srec(f0, spectral.ROLE_REFLECTANCE)
        # roughness defaults: ggx 0.5, rough_dielectric 0.2 (rough_dielectric.ts:17)
        rough = float(m.get("roughness", 0.2 if m["model"] == "rough_dielectric" else 0.5))
        rec["alpha"] = max(1e-3, rough * rough)                  # ggx.ts derived alpha
        tr = m.get("transmittance", [1.0, 1.0, 1.0])
        rec["transmittance"][:3] = spectral.rgb_of(tr)
        rec["s_transmittance"] = srec(tr, spectral.ROLE_REFLECTANCE)
        n_ref, kind, coef = index_record(m.get("ior", 1.5), spec, f"material '{n}': ior")
        rec["ior"] = n_ref
        rec["ior_kind"] = kind
        rec["ior_coef"][:] = coef
        med = m.get("medium")
        if med is not None:
            rec["has_medium"] = 1
            rec["medium_arm"] = medium_arm(med, strategy)
            # measurement.scattering 'ignored': the medium renders absorbing-only, shadows included
            sig_s = _rgb3(med.get("sigma_s", 0.0)) if strategy.scattering == "full" else [0.0, 0.0, 0.0]
            rec["sigma_a"][:3] = _rgb3(med.get("sigma_a", 0.0))
            rec["sigma_s"][:3] = sig_s
            rec["s_sigma_a"] = srec(_rgb3(med.get("sigma_a", 0.0)), spectral.ROLE_COEFFICIENT)
            rec["s_sigma_s"] = srec(sig_s, spectral.ROLE_COEFFICIENT)
            rec["hg_g"] = float(med.get("phase_g", 0.0))
            rec["phase"] = PHASE_MODELS[med.get("model", "hg")]
            eps = med.get("emission", 0.0)
            rec["medium_emission"][:3] = spectral.rgb_of(eps)
            rec["s_medium_emission"] = srec(eps, spectral.ROLE_EMISSION)
        materials.append(rec)

    tri_chunks, tri_obj_chunks, nrm_chunks, objects = [], [], [], []
    n_tris = 0

    def corner_normals(o, n):
        """Per-corner shading normals of a smooth mesh ([n][3][3], like the triangles), or None
        (flat: the geometric normal, MeshObject.normals absent, compiler/types.ts:140-141)."""
        if o.get("normals") is None:
            return None
        nr = np.asarray(o["normals"], np.float64).reshape(-1, 9).astype(np.float32)
        if nr.shape[0] != n:
            raise CompileError("mesh normals must give one normal per triangle corner")
        return nr

    def add_object(kind, material, closed=0, tris=None, p=None, normals=None):
        nonlocal n_tris
        rec = np.zeros(1, abi.OBJECT_DTYPE)[0]
        rec["kind"] = kind
        rec["material"] = material
        rec["closed"] = int(closed)
        if tris is not None:
            t32 = np.asarray(tris, np.float64).reshape(-1, 9).astype(np.float32)
            rec["first_tri"] = n_tris
            rec["tri_count"] = t32.shape[0]
            v = t32.reshape(-1, 3)
            rec["box_min"][:3] = v.min(axis=0)
            rec["box_max"][:3] = v.max(axis=0)
            tri_chunks.append(t32)
            tri_obj_chunks.append(np.full(t32.shape[0], len(objects), np.int32))
            nrm_chunks.append(normals if normals is not None else np.zeros_like(t32))
            rec["smooth"] = int(normals is not None)
            n_tris += t32.shape[0]
        if p is not None:
            rec["p"][:len(p)] = p
        objects.append(rec)
        return len(objects) - 1

    proto_chunks, proto_nrm_chunks, inst_recs = [], [], []
    n_proto = 0

    def add_batch(o, material):
        """An instance batch (features/intersection.ts:981-1066): prototype triangles in their
        local frame and placements {rotation: unit quaternion x y z w (local -> world),
        translation, scale}; the host precomposes each placement's inverse in f64 as the
        reference's driven-placement contract does (glsl/core/placement.glsl:1-17)."""
        nonlocal n_proto
        t32 = np.asarray(o["triangles"], np.float64).reshape(-1, 9).astype(np.float32)
        rec = np.zeros(1, abi.OBJECT_DTYPE)[0]
        rec["kind"] = abi.OBJ_INSTANCES
        rec["material"] = material
        rec["closed"] = int(o.get("closed", False))
        rec["first_tri"] = n_proto
        rec["tri_count"] = t32.shape[0]
        v = t32.reshape(-1, 3)
        rec["box_min"][:3] = v.min(axis=0)
        rec["box_max"][:3] = v.max(axis=0)
        rec["first_instance"] = len(inst_recs)
        proto_chunks.append(t32)
        pn = corner_normals(o, t32.shape[0])
        proto_nrm_chunks.append(pn if pn is not None else np.zeros_like(t32))
        rec["smooth"] = int(pn is not None)
        n_proto += t32.shape[0]
        q, t, sc = placement_arrays(o["placements"])
        if not np.all(sc > 0.0):
            raise CompileError("instance scale must be > 0")
        qq = q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3]
        q = q / np.sqrt(qq)[:, None]
        u, w = -q[:, :3], q[:, 3:4]                                    # world -> local: conj(q)
        rt = t + 2.0 * np.cross(u, np.cross(u, t) + w * t)             # R^-1 t
        ir = np.zeros(q.shape[0], abi.INSTANCE_DTYPE)
        ir["q"][:, :3] = u
        ir["q"][:, 3] = q[:, 3]
        ir["ts"][:, :3] = -rt
        ir["ts"][:, 3] = sc
        inst_recs.extend(ir)
        rec["instance_count"] = q.shape[0]
        objects.append(rec)

    for o in scene.get("objects", []):
        mi = mat_index[o["material"]]
        o = indexed_mesh_object(o)
        if o["type"] == "instances":
            add_batch(o, mi)
            continue
        if o["type"] == "mesh":
            tris = np.asarray(o["triangles"], np.float64).reshape(-1, 9)
            add_object(abi.OBJ_MESH, mi, o.get("closed", False), tris=tris, normals=corner_normals(o, tris.shape[0]))
        elif o["type"] == "sphere":
            add_object(abi.OBJ_SPHERE, mi, 1, p=list(o["center"]) + [o["radius"]])
        elif o["type"] == "quad":
            e1, e2 = np.asarray(o["edge1"], np.float64), np.asarray(o["edge2"], np.float64)
            cr = np.cross(e1, e2)
            add_object(abi.OBJ_QUAD, mi, 0, p=list(o["corner"]) + list(e1) + list(e2) + list(cr / np.linalg.norm(cr)))

    # Lights: an area light is a thin emissive quad region plus a light record
    # (lights/quad/quad.ts: region + derived normal/area; lighting.ts power CDF).
    lights = []
    light_tree_in = []     # per light: its tree box (6 f64) and selection power (sceneData.ts:99-143)
    for l in scene.get("lights", []):
        rec = np.zeros(1, abi.LIGHT_DTYPE)[0]
        rad_value = l.get("radiance", l.get("intensity", [1, 1, 1]))
        rad = np.asarray(spectral.rgb_of(rad_value), np.float64)
        rec["radiance"][:3] = rad
        rec["s_radiance"] = srec(rad_value, spectral.ROLE_EMISSION)
        if l["type"] == "quad":
            e1, e2 = np.asarray(l["edge1"], np.float64), np.asarray(l["edge2"], np.float64)
            cr = np.cross(e1, e2)
            area = float(np.linalg.norm(cr))
            nrm = cr / area
            em = np.zeros(1, abi.MATERIAL_DTYPE)[0]
            em["model"] = abi.MAT_LAMBERT
            em["emission"][:3] = rad
            em["s_emission"] = srec(rad_value, spectral.ROLE_EMISSION)
            em["emissive"] = 1
            materials.append(em)
            region = add_object(abi.OBJ_QUAD, len(materials) - 1, 0,
                                p=list(l["corner"]) + list(e1) + list(e2) + list(nrm))
            rec["kind"] = abi.LIGHT_QUAD
            rec["region"] = region
            rec["corner"][:3] = l["corner"]
            rec["edge1"][:3] = e1
            rec["edge2"][:3] = e2
            rec["normal"][:3] = nrm
            rec["area"] = area
            power = max(1e-8, math.pi * area * ((float(rad[0]) + float(rad[1])) + float(rad[2])) / 3.0)
            # treeBounds (lights/quad/quad.ts:66-78): the four corners
            c0 = np.asarray(l["corner"], np.float64)
            corners = np.array([c0 + s1 * e1 + s2 * e2 for s1, s2 in ((0, 0), (1, 0), (0, 1), (1, 1))])
            tree_box = list(corners.min(0)) + list(corners.max(0))
        else:
            rec["kind"] = abi.LIGHT_POINT
            rec["region"] = -1
            rec["corner"][:3] = l["position"]
            power = max(1e-8, 4.0 * math.pi * ((float(rad[0]) + float(rad[1])) + float(rad[2])) / 3.0)
            tree_box = [float(v) for v in l["position"]] * 2       # treeBounds: the position (point.ts:30-33)
        lights.append((rec, power))
        light_tree_in.append((tree_box, power))
    light_recs = []
    if lights:
        pdfs, cdf = select_pdfs([p for _, p in lights], strategy.light_selection)
        for (rec, _), p, c in zip(lights, pdfs, cdf):
            rec["select_pdf"] = p
            rec["cdf"] = c
            light_recs.append(rec)

    env = abi.PtEnv(abi.ENV_NONE, 512, 256, 0, 1.0, 0.0, 0.5, 0)
    e = scene.get("environment", {"type": "none"})
    if e.get("type") == "procedural":
        # env.size (u_envSize): the importance table's W x H, 512 x 256 unless a registered table's
        tw, th = e.get("table_size", (512, 256))
        env = abi.PtEnv(abi.ENV_SKY, int(tw), int(th), int(e.get("compensated", False)),
                        float(e.get("intensity", 1.0)), float(e.get("rotation", 0.0)),
                        float(e.get("select_prob", 0.5)), 1)
        env.body = sky_body(e)
        # estimator.envSampler.chart (components/env/index.ts ENV_CHARTS): the table's chart
        chart = e.get("chart", "equirect")
        if chart not in ENV_CHARTS:
            raise CompileError(f"environment.chart {chart!r}: expected one of {sorted(ENV_CHARTS)}")
        env.chart = ENV_CHARTS[chart]
    elif e.get("type") == "image":
        # features/environment.ts:144-190: the equirect picture; its table on the sampler's chart
        # (the picture's own W x H, or N x N octahedral, N = table_size[1], 256 default)
        img = np.ascontiguousarray(e["image"], np.float32)
        tw, th = e.get("table_size", (512, 256))
        env = abi.PtEnv(abi.ENV_IMAGE, int(img.shape[1]), int(th), int(e.get("compensated", False)),
                        float(e.get("intensity", 1.0)), float(e.get("rotation", 0.0)),
                        float(e.get("select_prob", 0.5)), 1)
        chart = e.get("chart", "equirect")
        if chart not in ENV_CHARTS:
            raise CompileError(f"environment.chart {chart!r}: expected one of {sorted(ENV_CHARTS)}")
        env.chart = ENV_CHARTS[chart]
        if env.chart == abi.CHART_EQUIRECT:
            env.table_w, env.table_h = int(img.shape[1]), int(img.shape[0])
    elif e.get("type") == "constant":
        # features/environment.ts:119-133: a constant sky, sampled as a light only when asked
        col = e.get("color", [1.0, 1.0, 1.0])
        env = abi.PtEnv(abi.ENV_CONSTANT, 0, 0, 0, float(e.get("intensity", 1.0)), 0.0,
                        float(e.get("select_prob", 0.5)), int(bool(e.get("sample_as_light", False))))
        for i, c in enumerate(spectral.rgb_of(col)):
            env.color[i] = float(c)
        env.s_color = abi.PtSpectrum.of(srec(col, spectral.ROLE_EMISSION))

    if env.type != abi.ENV_NONE and (env.type in (abi.ENV_SKY, abi.ENV_IMAGE) or env.samplable) and light_recs \
            and "select_prob" not in e:
        total = e.get("total_weight")
        if env.type == abi.ENV_IMAGE and total is None:
            from . import _lib
            img = np.ascontiguousarray(e["image"], np.float32)
            total = _lib.load().pt_env_image_total_weight(env, img.ctypes.data, img.shape[1], img.shape[0])
        env.select_prob = env_select_probability(scene, env, [p for _, p in lights], strategy.light_selection, total)

    tv = np.concatenate(tri_chunks) if tri_chunks else np.zeros((0, 9), np.float32)
    to = np.concatenate(tri_obj_chunks) if tri_obj_chunks else np.zeros(0, np.int32)
    amb = scene.get("ambient_medium")
    if amb is not None and amb not in mat_index:
        raise CompileError(f"ambientMedium references unknown material '{amb}'")
    # a spectral program lifts a coloured sky body per direction through the coefficient atlas
    atlas = spec and ((env.type == abi.ENV_SKY and env.body not in GREY_SKY_BODIES) or env.type == abi.ENV_IMAGE)
    arrays = abi.SceneArrays(tv, to, np.array(objects, abi.OBJECT_DTYPE),
                             np.array(materials, abi.MATERIAL_DTYPE),
                             np.array(light_recs, abi.LIGHT_DTYPE), env,
                             mat_index[amb] if amb is not None else -1,
                             spectral.curve_rows(), spectral.XYZ_TO_RGB.reshape(9), spectral.CIE_Y_INTEGRAL,
                             np.concatenate(proto_chunks) if proto_chunks else None,
                             np.array(inst_recs, abi.INSTANCE_DTYPE) if inst_recs else None,
                             np.concatenate(nrm_chunks) if any(r["smooth"] and r["kind"] == abi.OBJ_MESH for r in objects) else None,
                             np.concatenate(proto_nrm_chunks) if any(r["smooth"] and r["kind"] == abi.OBJ_INSTANCES
                                                                     for r in objects) else None,
                             spectral.atlas() if atlas else None, spectral.z_nodes() if atlas else None,
                             np.array([b for b, _ in light_tree_in], np.float64) if light_tree_in else None,
                             np.array([p for _, p in light_tree_in], np.float64) if light_tree_in else None,
                             np.ascontiguousarray(e["image"], np.float32) if e.get("type") == "image" else None)

    prog = abi.PtProgram()
    prog.abi_version = abi.ABI_VERSION
    prog.estimator = ESTIMATORS[strategy.direct_lighting]
    prog.max_bounces = strategy.max_bounces
    prog.rr_start = strategy.rr_start
    prog.rr_max_survival = strategy.rr_max_survival
    prog.color = abi.COLOR_SPECTRAL if spec else abi.COLOR_RGB
    prog.wavelengths = abi.WL_UNIFORM if strategy.wavelengths == "uniform" else abi.WL_VISIBLE
    prog.medium_light_sampling = abi.MLS_EQUIANGULAR if strategy.medium_light_sampling == "equiangular" else abi.MLS_VERTEX
    prog.volume_sampling = abi.VS_DELTA_TRACKING if strategy.volume_sampling == "delta-tracking" else abi.VS_ANALYTIC
    prog.variance = int(bool(strategy.variance))
    prog.accumulation = ACCUMULATIONS[strategy.accumulation]
    prog.aovs = sum({"albedo": abi.AOV_ALBEDO, "normal": abi.AOV_NORMAL}[a] for a in set(strategy.aovs))
    prog.env_chart = (ENV_CHARTS[e.get("chart", "equirect")] + 1 if e.get("type") in ("procedural", "image")
                      else abi.CHART_EQUIRECT + 1)
    prog.width, prog.height = width, height
    prog.frames_in_flight = max(1, strategy.frames_in_flight)
    prog.max_null_crossings = strategy.max_null_crossings
    prog.light_selection = {"power": abi.SEL_POWER, "uniform": abi.SEL_UNIFORM, "bvh": abi.SEL_BVH}[strategy.light_selection]
    warnings = []
    if strategy.volume_sampling == "delta-tracking" and not any(int(r["medium_arm"]) == abi.MED_DELTA and r["has_medium"]
                                                                 for r in materials):
        warnings.append("volumeSampling 'delta-tracking' controls nothing here — no medium is emissive and "
                        "scattering; constant media stay on the analytic arms")
    sel = "" if strategy.light_selection == "power" else "-" + strategy.light_selection
    return CompiledRenderer(id=f"{strategy.color}-{strategy.direct_lighting}{sel}-{scene.get('id', 'scene')}", scene=arrays,
                            program=prog, width=width, height=height, camera=scene["camera"],
                            reset_salt=reset_salt, warnings=warnings, material_names=list(mat_names),
                            source=scene, strategy=strategy)


# The procedural sky bodies (environment.glsl.source) the library evaluates: the reference's
# proc-sky (tests/compiler/__snapshots__/generated-glsl.snapshot.test.ts.snap:63758),
# proc-sky-rotated (tests/witnesses/scenes/estimatorAgreementWitness.ts:61-72), sky-missed-cap and
# spectral-comp-empty (tests/witnesses/scenes/spectralSkies.ts:39-48).
PROC_SKY_GLSL = ("mix(vec3(0.06, 0.08, 0.12), vec3(0.35, 0.45, 0.65), 0.5 * (dir.y + 1.0)) + vec3(18.0, 16.0, 13.0) * "
                 "smoothstep(0.9994, 0.9999, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))")
BIG_SUN_GLSL = "vec3(0.02) + vec3(40.0) * smoothstep(0.985, 0.99, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))"
CAP_GLSL = "vec3(step(0.9, dir.y))"
WHITE_GLSL = "vec3(1.0)"
SKY_BODY_GLSL = {abi.SKY_SUN: PROC_SKY_GLSL, abi.SKY_BIG_SUN: BIG_SUN_GLSL, abi.SKY_CAP: CAP_GLSL,
                 abi.SKY_WHITE: WHITE_GLSL}
GREY_SKY_BODIES = (abi.SKY_BIG_SUN, abi.SKY_CAP, abi.SKY_WHITE)


def _squash(src):
    return "".join(src.split())


def sky_body(env):
    """The PT_SKY_* body of a procedural environment: its 'glsl' source (a string or the
    reference's {'kind': 'glsl', 'source': ...}), whitespace aside, must be one of the bodies
    above; no 'glsl' is the proc-sky body."""
    g = env.get("glsl")
    if g is None:
        return abi.SKY_SUN
    src = g.get("source") if isinstance(g, dict) else g
    if not isinstance(src, str):
        raise CompileError("environment.glsl: expected a GLSL source string")
    for body, known in SKY_BODY_GLSL.items():
        if _squash(src) == _squash(known):
            return body
    raise CompileError(f"environment 'procedural' with the GLSL body {src!r} is not supported by this backend "
                       f"(known bodies: {', '.join(repr(v) for v in SKY_BODY_GLSL.values())})")


def program_from_description(description, width, height, frames_in_flight=1):
    """The library's lowering of a reference ProgramDescription (a dict, or its JSON text) to a
    pt_program (pt_program_from_description); raises _lib.PtError naming the unsupported field."""
    import json as _json
    from . import _lib
    text = description if isinstance(description, str) else _json.dumps(description)
    raw = text.encode()
    prog = abi.PtProgram()
    lib = _lib.load()
    _lib.check(lib.pt_program_from_description(raw, len(raw), int(width), int(height), int(frames_in_flight), C_byref(prog)))
    return prog


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


def describe(scene, strategy):
    """The ProgramDescription fields (src/compiler/plan/types.ts:39-249) this backend's plan
    decides for (scene, strategy), in the reference's vocabulary: what the reference's Planner
    would record for the same choices. Round trip: program_from_description(describe(...))
    equals compile_scene(...).program for the lowered fields."""
    lights = scene.get("lights", [])
    env = scene.get("environment", {"type": "none"})
    models = []
    for m in scene.get("materials", {}).values():
        if m["model"] not in models:
            models.append(m["model"])
    media_models = []
    for m in scene.get("materials", {}).values():
        med = m.get("medium")
        if med is not None and medium_arm(med, strategy) != abi.MED_ABSORB and med.get("model", "hg") not in media_models:
            media_models.append(med.get("model", "hg"))
    # the spectral kinds the program evaluates (a Gaussian band is a reflectance of a material)
    kinds = []
    if strategy.color == "spectral":
        for m in scene.get("materials", {}).values():
            for k in ("albedo", "f0", "transmittance"):
                if spectral.is_gaussian(m.get(k)) and "gaussian" not in kinds:
                    kinds.append("gaussian")
    d = {
        "dispersion": None,
        "environment": ({"type": "procedural", "glsl": {"kind": "glsl", "source": SKY_BODY_GLSL[sky_body(env)]}}
                        if env.get("type") == "procedural" else {"type": env.get("type", "none")}),
        "estimator": {
            "accumulation": {"type": strategy.accumulation, **({"variance": True} if strategy.variance else {})},
            "envSampler": {"chart": env.get("chart", "equirect"), "compensation": bool(env.get("compensated", False))},
            "lighting": None if strategy.direct_lighting == "none" else {"method": strategy.direct_lighting,
                                                                          "selection": strategy.light_selection},
            "mediumLightSampling": strategy.medium_light_sampling,
            "russianRoulette": None if strategy.rr_start < 0 else {"startDepth": strategy.rr_start,
                                                                    "maxSurvival": strategy.rr_max_survival},
            "volumeSampling": strategy.volume_sampling if media_models else "none",
            "wavelengths": strategy.wavelengths if strategy.color == "spectral" else None,
        },
        "intersection": {"drivenPlacement": False, "instanceAccel": "tlas", "meshTraversal": "bvh",
                         "objectDispatch": "unrolled"},
        "materials": {"models": models, "twoSidedShading": False, "materialsReadUv": False},
        "measurement": {
            "ambient": "euclidean",
            "camera": {"type": scene["camera"].get("type", "pinhole"),
                       **{k: v for k, v in scene["camera"].items() if k != "type"}},
            "color": strategy.color, "maxBounces": strategy.max_bounces,
            "maxNullCrossings": strategy.max_null_crossings, "response": "radiance",
            "scattering": strategy.scattering, "sensor": "cie1931" if strategy.color == "spectral" else "ideal",
            "shadows": "opaque-dielectrics",
            **({"aovs": list(strategy.aovs)} if strategy.aovs else {}),
        },
        "media": {"deflecting": False, "heterogeneousArms": False, "models": media_models},
        "spectrumKinds": kinds,
        "view": {"tonemap": {"type": "reinhard"}},
    }
    _ = lights
    return d
