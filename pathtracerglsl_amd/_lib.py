"""ctypes binding of libptmi.so (the in-tree HIP build). Fails loudly when it is missing:
there is no CPU fallback for any entry point."""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptmi.so")

# every symbol declared in include/ptmi.h (the boundary)
DECLARED = ["pt_last_error", "pt_ctx_create", "pt_ctx_destroy", "pt_scene_upload", "pt_renderer_load",
            "pt_renderer_unload", "pt_render_frames", "pt_read_export", "pt_clear_accumulation",
            "pt_sample_count", "pt_trace_closest", "pt_trace_any", "pt_trace_closest_device",
            "pt_trace_any_device", "pt_set_timing", "pt_get_timings", "pt_last_ray_counts", "pt_finish",
            "pt_scene_update_materials", "pt_forward", "pt_backward",
            "pt_batch_create", "pt_batch_destroy", "pt_batch_set_config", "pt_batch_get_config",
            "pt_batch_render", "pt_batch_read_export", "pt_batch_autotune",
            "pt_program_from_description", "pt_renderer_load_description", "pt_env_total_weight", "pt_env_image_total_weight",
            "pt_obj_parse", "pt_mesh_data_info", "pt_mesh_data_copy", "pt_mesh_data_free", "pt_inst_parse",
            "pt_render_frames_noalloc", "pt_bvh_build", "pt_light_tree_build", "pt_mesh_bvh_build", "pt_plan_data_layout",
            "pt_pack_mesh_channels", "pt_ctx_reset", "pt_scene_update_lights", "pt_scene_update_env",
            "pt_env_table_upload"]
# every symbol declared in include/ptmi_debug.h (test and bench instrumentation)
DECLARED_DEBUG = ["pt_renderer_set_traversal_counting", "pt_last_traversal_counts", "pt_trace_closest_device_counted",
                  "pt_scene_stats", "pt_abi_sizeof", "pt_debug_wide_bvh", "pt_debug_math", "pt_calib_copy",
                  "pt_debug_env_bake", "pt_debug_env_table", "pt_debug_region_at", "pt_debug_env_sample"]

_lib = None


# pt_status codes (include/ptmi.h)
E_INVALID, E_MISSING_INPUT, E_DEVICE, E_OOM, E_UNSUPPORTED, E_CONTEXT_LOST = 1, 2, 3, 4, 5, 6


class PtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ptmi status {status}: {msg}")
        self.status = status


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libptmi.so not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, i32, u32, st = C.c_void_p, C.c_int32, C.c_uint32, C.c_int
    lib.pt_last_error.restype = C.c_char_p
    lib.pt_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.pt_ctx_destroy.argtypes = [vp]
    lib.pt_ctx_destroy.restype = None
    lib.pt_finish.argtypes = [vp]
    lib.pt_scene_upload.argtypes = [vp, C.POINTER(abi.PtSceneDesc)]
    lib.pt_renderer_load.argtypes = [vp, C.POINTER(abi.PtProgram), C.POINTER(vp)]
    lib.pt_renderer_unload.argtypes = [vp]
    lib.pt_renderer_unload.restype = None
    lib.pt_render_frames.argtypes = [vp, C.POINTER(abi.PtParams), u32]
    lib.pt_render_frames_noalloc.argtypes = [vp, C.POINTER(abi.PtParams), u32]
    lib.pt_read_export.argtypes = [vp, C.c_char_p, vp, C.c_size_t]
    lib.pt_clear_accumulation.argtypes = [vp]
    lib.pt_sample_count.argtypes = [vp]
    lib.pt_sample_count.restype = u32
    lib.pt_trace_closest.argtypes = [vp, C.POINTER(abi.PtRaysSoa), u32, vp, vp]
    lib.pt_trace_any.argtypes = [vp, C.POINTER(abi.PtRaysSoa), u32, vp]
    lib.pt_trace_closest_device.argtypes = [vp, vp, u32, vp, vp]
    lib.pt_trace_any_device.argtypes = [vp, vp, u32, vp]
    lib.pt_set_timing.argtypes = [vp, C.c_int]
    lib.pt_get_timings.argtypes = [vp, vp, C.c_int]
    lib.pt_get_timings.restype = C.c_int
    lib.pt_last_ray_counts.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.pt_renderer_set_traversal_counting.argtypes = [vp, C.c_int]
    lib.pt_last_traversal_counts.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.pt_trace_closest_device_counted.argtypes = [vp, vp, u32, vp, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.pt_scene_stats.argtypes = [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    lib.pt_scene_update_materials.argtypes = [vp, vp, i32]
    lib.pt_scene_update_lights.argtypes = [vp, vp, i32]
    lib.pt_scene_update_env.argtypes = [vp, C.POINTER(abi.PtEnv)]
    lib.pt_env_table_upload.argtypes = [vp, vp, i32, i32, i32, i32, C.POINTER(C.c_double)]
    lib.pt_ctx_reset.argtypes = [vp]
    lib.pt_debug_region_at.argtypes = [vp, vp, C.c_int, vp]
    lib.pt_debug_env_sample.argtypes = [vp, vp, C.c_int, vp]
    lib.pt_forward.argtypes = [vp, C.POINTER(abi.PtParams), u32]
    lib.pt_backward.argtypes = [vp, C.POINTER(abi.PtParams), u32, C.POINTER(abi.PtAdjointParams), vp, vp]
    lib.pt_batch_create.argtypes = [vp, i32, vp, vp, vp, C.POINTER(vp)]
    lib.pt_batch_destroy.argtypes = [vp]
    lib.pt_batch_destroy.restype = None
    lib.pt_batch_set_config.argtypes = [vp, C.POINTER(abi.PtBatchConfig)]
    lib.pt_batch_get_config.argtypes = [vp, C.POINTER(abi.PtBatchConfig)]
    lib.pt_batch_render.argtypes = [vp, vp]
    lib.pt_batch_read_export.argtypes = [vp, i32, vp, C.c_size_t]
    lib.pt_batch_autotune.argtypes = [vp, vp, u32, C.POINTER(abi.PtBatchConfig), vp, vp, i32, C.POINTER(i32)]
    lib.pt_program_from_description.argtypes = [C.c_char_p, C.c_size_t, i32, i32, i32, C.POINTER(abi.PtProgram)]
    lib.pt_renderer_load_description.argtypes = [vp, C.c_char_p, C.c_size_t, i32, i32, i32, C.POINTER(vp)]
    lib.pt_env_total_weight.argtypes = [C.POINTER(abi.PtEnv)]
    lib.pt_env_total_weight.restype = C.c_double
    lib.pt_env_image_total_weight.argtypes = [C.POINTER(abi.PtEnv), vp, i32, i32]
    lib.pt_env_image_total_weight.restype = C.c_double
    lib.pt_obj_parse.argtypes = [C.c_char_p, C.c_size_t, u32, C.POINTER(vp)]
    lib.pt_mesh_data_info.argtypes = [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(i32), C.POINTER(i32)]
    lib.pt_mesh_data_copy.argtypes = [vp, vp, vp, vp, vp]
    lib.pt_mesh_data_free.argtypes = [vp]
    lib.pt_mesh_data_free.restype = None
    lib.pt_inst_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(abi.PtInstTable)]
    lib.pt_bvh_build.argtypes = [vp, i32, i32, u32, vp, vp, C.POINTER(i32), C.POINTER(i32)]
    lib.pt_light_tree_build.argtypes = [vp, vp, i32, vp, vp, C.POINTER(i32), C.POINTER(i32)]
    lib.pt_mesh_bvh_build.argtypes = [vp, i32, vp, i32, vp, vp, C.POINTER(i32), C.POINTER(i32)]
    lib.pt_plan_data_layout.argtypes = [vp, vp]
    lib.pt_pack_mesh_channels.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp]
    lib.pt_abi_sizeof.argtypes = [C.c_int]
    lib.pt_abi_sizeof.restype = C.c_size_t
    _ = (i32, st)
    _lib = lib
    return lib


def check(status):
    if status != 0:
        raise PtError(status, load().pt_last_error().decode(errors="replace"))
