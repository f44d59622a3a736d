"""ctypes / numpy mirrors of the C ABI records declared in include/ptmi.h.

Field order and sizes must match the header exactly; tests/test_abi.py checks the sizes
against the compiled library's own sizeof() export.
"""
import ctypes as C

import numpy as np

ABI_VERSION = 15

OBJ_MESH, OBJ_QUAD, OBJ_SPHERE, OBJ_INSTANCES = 0, 1, 2, 3
MAT_LAMBERT, MAT_DIELECTRIC, MAT_GGX, MAT_NULL, MAT_MIRROR, MAT_ROUGH_DIELECTRIC, MAT_CHECKER = 0, 1, 2, 3, 4, 5, 6
MED_ANALYTIC, MED_ABSORB, MED_DELTA = 0, 1, 2
VS_ANALYTIC, VS_DELTA_TRACKING = 0, 1
IOR_CONSTANT, IOR_CAUCHY, IOR_SELLMEIER = 0, 1, 2
LIGHT_QUAD, LIGHT_POINT = 0, 1
ENV_NONE, ENV_SKY, ENV_CONSTANT, ENV_IMAGE = 0, 1, 2, 3
SKY_SUN, SKY_BIG_SUN, SKY_CAP, SKY_WHITE = 0, 1, 2, 3
AOV_ALBEDO, AOV_NORMAL = 1, 2
CHART_EQUIRECT, CHART_OCTAHEDRAL = 0, 1
SEL_POWER, SEL_UNIFORM, SEL_BVH = 0, 1, 2
CAM_PINHOLE, CAM_ORTHOGRAPHIC, CAM_THINLENS, CAM_FISHEYE, CAM_CYLINDRICAL, CAM_EQUIRECT = 0, 1, 2, 3, 4, 5
FISHEYE_PROJECTIONS = {"equidistant": 0, "equisolid": 1, "stereographic": 2, "orthographic": 3}
EST_PT, EST_NEE, EST_MIS = 0, 1, 2
COLOR_RGB, COLOR_SPECTRAL = 0, 1
WL_VISIBLE, WL_UNIFORM = 0, 1
MLS_VERTEX, MLS_EQUIANGULAR = 0, 1
SPEC_ZERO, SPEC_GREY, SPEC_SIGMOID, SPEC_BLACKBODY, SPEC_GAUSSIAN = 0, 1, 2, 3, 4
PHASE_HG, PHASE_RAYLEIGH = 0, 1
ACC_AVERAGE, ACC_ONESHOT = 0, 1

SPECTRUM_DTYPE = np.dtype([
    ("kind", "<i4"), ("scaled", "<i2"), ("d65", "<i2"), ("value", "<f4"), ("scale", "<f4"),
    ("c", "<f4", 3), ("pad", "<f4"),
])

OBJECT_DTYPE = np.dtype([
    ("kind", "<i4"), ("material", "<i4"), ("closed", "<i4"), ("first_tri", "<i4"),
    ("tri_count", "<i4"), ("first_instance", "<i4"), ("instance_count", "<i4"), ("smooth", "<i4"),
    ("p", "<f4", 12), ("box_min", "<f4", 4), ("box_max", "<f4", 4),
])
MATERIAL_DTYPE = np.dtype([
    ("model", "<i4"), ("emissive", "<i4"), ("has_medium", "<i4"), ("medium_arm", "<i4"),
    ("albedo", "<f4", 4), ("emission", "<f4", 4), ("f0", "<f4", 4), ("transmittance", "<f4", 4),
    ("alpha", "<f4"), ("ior", "<f4"), ("hg_g", "<f4"), ("pad1", "<f4"),
    ("sigma_s", "<f4", 4), ("sigma_a", "<f4", 4),
    ("s_albedo", SPECTRUM_DTYPE), ("s_emission", SPECTRUM_DTYPE), ("s_f0", SPECTRUM_DTYPE),
    ("s_transmittance", SPECTRUM_DTYPE), ("s_sigma_s", SPECTRUM_DTYPE), ("s_sigma_a", SPECTRUM_DTYPE),
    ("medium_emission", "<f4", 4), ("s_medium_emission", SPECTRUM_DTYPE), ("majorant", "<f4", 2),
    ("ior_kind", "<i4"), ("phase", "<i4"), ("ior_coef", "<f4", 8),
])
LIGHT_DTYPE = np.dtype([
    ("kind", "<i4"), ("region", "<i4"), ("select_pdf", "<f4"), ("cdf", "<f4"),
    ("corner", "<f4", 4), ("edge1", "<f4", 4), ("edge2", "<f4", 4), ("normal", "<f4", 4),
    ("radiance", "<f4", 4), ("area", "<f4"), ("pad", "<f4", 3), ("s_radiance", SPECTRUM_DTYPE),
])
INSTANCE_DTYPE = np.dtype([("q", "<f4", 4), ("ts", "<f4", 4)])
assert SPECTRUM_DTYPE.itemsize == 32 and INSTANCE_DTYPE.itemsize == 32
assert OBJECT_DTYPE.itemsize == 112 and MATERIAL_DTYPE.itemsize == 416 and LIGHT_DTYPE.itemsize == 144


class PtSpectrum(C.Structure):
    _fields_ = [("kind", C.c_int32), ("scaled", C.c_int16), ("d65", C.c_int16), ("value", C.c_float),
                ("scale", C.c_float), ("c", C.c_float * 3), ("pad", C.c_float)]

    @classmethod
    def of(cls, rec):
        s = cls()
        s.kind, s.scaled, s.d65 = int(rec["kind"]), int(rec["scaled"]), int(rec["d65"])
        s.value, s.scale = float(rec["value"]), float(rec["scale"])
        for i in range(3):
            s.c[i] = float(rec["c"][i])
        return s


class PtEnv(C.Structure):
    _fields_ = [("type", C.c_int32), ("table_w", C.c_int32), ("table_h", C.c_int32),
                ("compensated", C.c_int32), ("intensity", C.c_float), ("rotation", C.c_float),
                ("select_prob", C.c_float), ("samplable", C.c_int32), ("color", C.c_float * 4),
                ("s_color", PtSpectrum), ("chart", C.c_int32), ("body", C.c_int32)]


class PtSceneDesc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("n_tris", C.c_int32),
                ("tri_vertices", C.c_void_p), ("tri_object", C.c_void_p),
                ("n_objects", C.c_int32), ("objects", C.c_void_p),
                ("n_materials", C.c_int32), ("materials", C.c_void_p),
                ("n_lights", C.c_int32), ("lights", C.c_void_p),
                ("env", PtEnv), ("ambient_material", C.c_int32), ("pad0", C.c_int32),
                ("spectral_curves", C.c_void_p), ("xyz_to_rgb", C.c_float * 9),
                ("cie_y_integral", C.c_float), ("pad1", C.c_int32 * 2),
                ("n_proto_tris", C.c_int32), ("n_instances", C.c_int32),
                ("proto_vertices", C.c_void_p), ("instances", C.c_void_p),
                ("tri_normals", C.c_void_p), ("proto_normals", C.c_void_p),
                ("spectral_atlas", C.c_void_p), ("spectral_z_nodes", C.c_void_p),
                ("light_boxes", C.c_void_p), ("light_powers", C.c_void_p),
                ("env_image", C.c_void_p), ("env_image_w", C.c_int32), ("env_image_h", C.c_int32)]


class PtInstTable(C.Structure):
    """pt_inst_table: a parsed `.inst` header, byte offsets into the caller's buffer."""
    _fields_ = [("count", C.c_uint32), ("flags", C.c_uint32), ("n_scalars", C.c_uint32), ("pad", C.c_uint32),
                ("aabb", C.c_float * 6), ("names_offset", C.c_uint64), ("provenance_offset", C.c_uint64),
                ("provenance_len", C.c_uint64), ("positions_offset", C.c_uint64), ("sizes_offset", C.c_uint64),
                ("colors_offset", C.c_uint64), ("orientations_offset", C.c_uint64), ("scalars_offset", C.c_uint64)]


OBJ_SMOOTH_NORMALS = 1


class PtProgram(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("estimator", C.c_int32), ("max_bounces", C.c_int32),
                ("rr_start", C.c_int32), ("rr_max_survival", C.c_float), ("color", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("frames_in_flight", C.c_int32),
                ("max_null_crossings", C.c_int32), ("wavelengths", C.c_int32), ("medium_light_sampling", C.c_int32),
                ("variance", C.c_int32), ("volume_sampling", C.c_int32), ("aovs", C.c_int32),
                ("env_chart", C.c_int32), ("light_selection", C.c_int32), ("accumulation", C.c_int32)]


class PtParams(C.Structure):
    _fields_ = [("cam_position", C.c_float * 4), ("cam_forward", C.c_float * 4),
                ("cam_right", C.c_float * 4), ("cam_up", C.c_float * 4),
                ("tan_fov", C.c_float), ("aspect", C.c_float), ("reset_salt", C.c_uint32),
                ("camera", C.c_int32), ("pixel_offset", C.c_int32 * 2), ("ortho_scale", C.c_float),
                ("aperture", C.c_float), ("focus_distance", C.c_float), ("projection", C.c_int32),
                ("fisheye_k", C.c_float), ("cyl_focal", C.c_float)]


ADJ_MAX_ALBEDO = 16
ADJ_N_PARAMS = 7 + ADJ_MAX_ALBEDO
REDUCE_ATOMIC, REDUCE_DETERMINISTIC = 0, 1


class PtAdjointParams(C.Structure):
    _fields_ = [("medium_material", C.c_int32), ("n_albedo", C.c_int32),
                ("albedo_material", C.c_int32 * ADJ_MAX_ALBEDO), ("reduce_mode", C.c_int32), ("pad", C.c_int32)]

    @classmethod
    def make(cls, medium_material, albedo_materials, reduce_mode=REDUCE_DETERMINISTIC):
        a = cls()
        a.medium_material = int(medium_material)
        a.n_albedo = len(albedo_materials)
        for k, m in enumerate(albedo_materials):
            a.albedo_material[k] = int(m)
        a.reduce_mode = int(reduce_mode)
        return a


class PtBatchConfig(C.Structure):
    _fields_ = [("wave_paths", C.c_int32), ("trace_block", C.c_int32), ("trace_blocks_per_cu", C.c_int32),
                ("shade_blocks_per_cu", C.c_int32), ("stack", C.c_int32), ("use_graph", C.c_int32),
                ("pad", C.c_int32 * 2)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "pad"}


class PtKernelTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_float), ("launches", C.c_uint32),
                ("items", C.c_uint64)]


class PtRaysSoa(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")]


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class SceneArrays:
    """Host-side scene in the ABI layout; keeps the numpy arrays alive for the desc."""

    def __init__(self, tri_vertices, tri_object, objects, materials, lights, env=None, ambient_material=-1,
                 spectral_curves=None, xyz_to_rgb=None, cie_y_integral=0.0, proto_vertices=None, instances=None,
                 tri_normals=None, proto_normals=None, spectral_atlas=None, spectral_z_nodes=None,
                 light_boxes=None, light_powers=None, env_image=None):
        self.tri_vertices = np.ascontiguousarray(tri_vertices, dtype=np.float32).reshape(-1, 9)
        self.tri_object = np.ascontiguousarray(tri_object, dtype=np.int32).reshape(-1)
        self.objects = np.ascontiguousarray(objects, dtype=OBJECT_DTYPE)
        self.materials = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        self.lights = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        self.env = env if env is not None else PtEnv(ENV_NONE, 512, 256, 0, 1.0, 0.0, 0.5, 0)
        self.ambient_material = int(ambient_material)
        self.spectral_curves = (None if spectral_curves is None
                                else np.ascontiguousarray(spectral_curves, dtype=np.float32).reshape(95, 4))
        self.xyz_to_rgb = np.zeros(9, np.float32) if xyz_to_rgb is None else np.asarray(xyz_to_rgb, np.float32).reshape(9)
        self.cie_y_integral = float(cie_y_integral)
        self.proto_vertices = (np.zeros((0, 9), np.float32) if proto_vertices is None
                               else np.ascontiguousarray(proto_vertices, dtype=np.float32).reshape(-1, 9))
        self.instances = (np.zeros(0, INSTANCE_DTYPE) if instances is None
                          else np.ascontiguousarray(instances, dtype=INSTANCE_DTYPE))
        self.tri_normals = None if tri_normals is None else np.ascontiguousarray(tri_normals, np.float32).reshape(-1, 9)
        self.proto_normals = None if proto_normals is None else np.ascontiguousarray(proto_normals, np.float32).reshape(-1, 9)
        self.spectral_atlas = (None if spectral_atlas is None
                               else np.ascontiguousarray(spectral_atlas, np.float32).reshape(3 * 64 ** 3, 4))
        self.spectral_z_nodes = (None if spectral_z_nodes is None
                                 else np.ascontiguousarray(spectral_z_nodes, np.float32).reshape(64))
        self.light_boxes = None if light_boxes is None else np.ascontiguousarray(light_boxes, np.float64).reshape(-1, 6)
        self.light_powers = None if light_powers is None else np.ascontiguousarray(light_powers, np.float64).reshape(-1)
        self.env_image = None if env_image is None else np.ascontiguousarray(env_image, np.float32)
        assert self.env_image is None or (self.env_image.ndim == 3 and self.env_image.shape[2] == 3)
        assert self.tri_object.shape[0] == self.tri_vertices.shape[0]
        assert self.tri_normals is None or self.tri_normals.shape == self.tri_vertices.shape
        assert self.proto_normals is None or self.proto_normals.shape == self.proto_vertices.shape

    @property
    def n_tris(self):
        return int(self.tri_vertices.shape[0])

    def desc(self):
        d = PtSceneDesc()
        d.abi_version = ABI_VERSION
        d.n_tris = self.n_tris
        d.tri_vertices = _ptr(self.tri_vertices)
        d.tri_object = _ptr(self.tri_object)
        d.n_objects = len(self.objects)
        d.objects = _ptr(self.objects)
        d.n_materials = len(self.materials)
        d.materials = _ptr(self.materials)
        d.n_lights = len(self.lights)
        d.lights = _ptr(self.lights)
        d.env = self.env
        d.ambient_material = self.ambient_material
        d.spectral_curves = _ptr(self.spectral_curves)
        for i in range(9):
            d.xyz_to_rgb[i] = float(self.xyz_to_rgb[i])
        d.cie_y_integral = self.cie_y_integral
        d.n_proto_tris = int(self.proto_vertices.shape[0])
        d.n_instances = int(self.instances.shape[0])
        d.proto_vertices = _ptr(self.proto_vertices)
        d.instances = _ptr(self.instances)
        d.tri_normals = _ptr(self.tri_normals)
        d.proto_normals = _ptr(self.proto_normals)
        d.spectral_atlas = _ptr(self.spectral_atlas)
        d.spectral_z_nodes = _ptr(self.spectral_z_nodes)
        d.light_boxes = _ptr(self.light_boxes)
        d.light_powers = _ptr(self.light_powers)
        d.env_image = _ptr(self.env_image)
        if self.env_image is not None:
            d.env_image_h, d.env_image_w = int(self.env_image.shape[0]), int(self.env_image.shape[1])
        return d


def rays_soa_array(o, d, tmax):
    """Pack rays into the 7 x n float32 block used by the trace entry points."""
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    n = o.shape[0]
    out = np.empty((7, n), np.float32)
    out[0:3] = o.T
    out[3:6] = d.T
    out[6] = np.broadcast_to(np.asarray(tmax, np.float32), (n,))
    return np.ascontiguousarray(out)
