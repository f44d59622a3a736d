"""ctypes loader for the CPU oracle (oracle/_build/liboracle.so) — test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

from pathtracerglsl_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(variant=""):
    path = os.path.join(ORACLE_DIR, "_build", f"liboracle{variant}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.or_scene_create.restype = vp
    lib.or_scene_create.argtypes = [C.POINTER(abi.PtSceneDesc)]
    lib.or_scene_destroy.argtypes = [vp]
    lib.or_trace_closest.argtypes = [vp, vp, C.c_int, vp, vp]
    lib.or_trace_closest_brute.argtypes = [vp, vp, C.c_int, vp, vp]
    lib.or_trace_any.argtypes = [vp, vp, C.c_int, vp]
    lib.or_render.argtypes = [vp, C.POINTER(abi.PtProgram), C.POINTER(abi.PtParams), C.c_uint32,
                              C.c_int, C.c_int, C.c_int, vp, vp, vp]
    lib.or_pcg4d.argtypes = [vp, vp, C.c_int]
    lib.or_random_stream.argtypes = [C.c_uint32] * 4 + [vp, C.c_int]
    lib.or_math_eval.argtypes = [C.c_int, vp, vp, vp, C.c_int]
    lib.or_env_table.argtypes = [vp, vp, vp]
    lib.or_scene_env_table_rgb.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.or_scene_env_table_rgb.restype = C.c_int
    lib.or_env_total.restype = C.c_double
    lib.or_env_total.argtypes = [vp]
    lib.or_region_at.argtypes = [vp, vp]
    lib.or_concentric_disk.argtypes = [vp, C.c_int, vp]
    lib.or_render_aovs.argtypes = [vp, vp, vp, C.c_uint32, C.c_int, vp, vp, vp, vp]
    lib.or_render_aovs.restype = None
    lib.or_concentric_disk.restype = None
    lib.or_oct_map.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, vp]
    lib.or_oct_map.restype = None
    lib.or_bsdf_probe.argtypes = [vp, vp, vp, C.c_float, C.c_float, C.c_float, vp]
    lib.or_env_pdf.argtypes = [vp, vp, C.c_int, vp]
    lib.or_env_sample.argtypes = [vp, vp, C.c_int, vp, vp]
    lib.or_phase_value.argtypes = [C.c_int, C.c_float, vp, C.c_int, vp]
    lib.or_phase_value.restype = None
    lib.or_phase_sample.argtypes = [C.c_int, C.c_float, vp, vp, C.c_int, vp, vp]
    lib.or_phase_sample.restype = None
    lib.or_spectrum_eval.argtypes = [vp, vp, vp, C.c_int, vp]
    lib.or_spectrum_eval.restype = None
    lib.or_env_radiance.argtypes = [vp, vp, vp, C.c_int, vp]
    lib.or_env_radiance.restype = None
    lib.or_resample_equirect_to_octahedral.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
    lib.or_resample_equirect_to_octahedral.restype = None
    lib.or_light_tree_build.argtypes = [vp, vp, C.c_int, vp, vp, C.POINTER(C.c_int)]
    lib.or_light_tree_build.restype = C.c_int
    lib.or_light_tree_importance.argtypes = [vp, vp, vp, vp, C.c_int]
    lib.or_light_tree_importance.restype = C.c_float
    lib.or_light_tree_pmf.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int]
    lib.or_light_tree_pmf.restype = C.c_float
    lib.or_light_tree_pick.argtypes = [vp, C.c_float, vp, vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float)]
    lib.or_light_tree_pick.restype = C.c_int
    lib.or_basic_eval.argtypes = [C.c_int, vp, vp, vp, C.c_int]
    lib.or_adjoint.argtypes = [vp, C.POINTER(abi.PtProgram), C.POINTER(abi.PtParams), C.c_int,
                               C.POINTER(abi.PtAdjointParams), vp, vp]
    lib.or_bvh_build.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    lib.or_bvh_build.restype = C.c_int
    lib.or_slab_probe.argtypes = [vp, vp, C.c_int, vp]
    lib.or_slab_probe.restype = None
    lib.or_set_proposal.argtypes = [C.c_int, vp, vp, C.c_float]
    lib.or_set_proposal.restype = None
    if variant == "_plants":
        lib.or_set_plant.argtypes = [C.c_int]
        lib.or_set_plant.restype = None
    return lib


class OracleScene:
    def __init__(self, lib, arrays):
        self.lib = lib
        self.arrays = arrays
        self._desc = arrays.desc()
        self.h = lib.or_scene_create(C.byref(self._desc))

    def close(self):
        if self.h:
            self.lib.or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def register_env_table(self, rgb, width, height, chart=0, compensated=False):
        """The scene's sky table rebuilt from an RGB picture (registerEnvironmentTable); returns
        the table's totalWeight."""
        a = np.ascontiguousarray(rgb, np.float32)
        tw = C.c_double()
        if self.lib.or_scene_env_table_rgb(self.h, a.ctypes.data, width, height, chart, int(bool(compensated)),
                                           C.byref(tw)) != 0:
            raise ValueError("or_scene_env_table_rgb rejected the picture")
        return tw.value

    def trace_closest(self, rays, brute=False):
        n = rays.shape[1]
        t = np.empty(n, np.float32)
        p = np.empty(n, np.int32)
        fn = self.lib.or_trace_closest_brute if brute else self.lib.or_trace_closest
        fn(self.h, rays.ctypes.data, n, t.ctypes.data, p.ctypes.data)
        return t, p

    def trace_any(self, rays):
        n = rays.shape[1]
        h = np.empty(n, np.uint8)
        self.lib.or_trace_any(self.h, rays.ctypes.data, n, h.ctypes.data)
        return h

    def render(self, program, params, nframes, sample_count=0, rows=None, keep_samples=False):
        W, H = program.width, program.height
        y0, y1 = rows if rows is not None else (0, H)
        film = np.zeros((H, W, 4), np.float32)
        samples = np.zeros((nframes, H, W, 3), np.float32) if keep_samples else None
        counts = np.zeros(2, np.uint64)
        self.lib.or_render(self.h, C.byref(program), C.byref(params), sample_count, nframes, y0, y1,
                           film.ctypes.data, samples.ctypes.data if samples is not None else None,
                           counts.ctypes.data)
        return film, samples, counts

    def render_aovs(self, program, params, nframes, sample_count=0):
        """Film and the two first-hit guide films (albedo, normal), each [H, W, 4]."""
        W, H = program.width, program.height
        film, alb, nrm = (np.zeros((H, W, 4), np.float32) for _ in range(3))
        counts = np.zeros(2, np.uint64)
        self.lib.or_render_aovs(self.h, C.byref(program), C.byref(params), sample_count, nframes, film.ctypes.data,
                                alb.ctypes.data, nrm.ctypes.data, counts.ctypes.data)
        return film, alb, nrm, counts

    def adjoint(self, program, params, nframes, adj, dL):
        """FP64 replay of the path-replay adjoint: d(sum dL . film)/dtheta."""
        dL = np.ascontiguousarray(dL, np.float32)
        assert dL.shape == (program.height, program.width, 4)
        g = np.zeros(abi.ADJ_N_PARAMS, np.float64)
        self.lib.or_adjoint(self.h, C.byref(program), C.byref(params), nframes, C.byref(adj), dL.ctypes.data, g.ctypes.data)
        return g

    def env_table(self, w=512, h=256):
        cond = np.zeros((h, w), np.float32)
        marg = np.zeros(h, np.float32)
        self.lib.or_env_table(self.h, cond.ctypes.data, marg.ctypes.data)
        return cond, marg

    def region_at(self, p):
        p = np.asarray(p, np.float32)
        return self.lib.or_region_at(self.h, p.ctypes.data)
