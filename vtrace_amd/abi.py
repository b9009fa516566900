"""ctypes declarations of include/vtrace.h (the C ABI of librender_hip).

This is the binding a Python host would add; the Rust host of the reference binds the
same seven symbols in ``src/render.rs:110-128``.
"""
from __future__ import annotations

import ctypes as C

# Symbols of the reference's FFI surface (src/render.rs:110-128) and the headless extensions.
REFERENCE_EXPORTS = ("entry", "render_tick", "get_input_data_pointer", "add_texture", "start_update_instances",
                     "end_update_instances", "cleanup")
EXTENSION_EXPORTS = ("vt_set_params", "vt_get_params", "vt_add_volume_indexed", "vt_add_textures", "vt_lbvh_builds",
                     "vt_set_cursor", "vt_set_frame_split", "vt_set_persistent_blocks", "vt_refit_instances",
                     "vt_read_hits", "vt_read_radiance",
                     "vt_read_bgra8", "vt_read_stats", "vt_read_stage_stats", "vt_last_frame_ms", "vt_bench_frames",
                     "vt_bench_stages", "vt_read_volume", "vt_read_brick_map",
                     "vt_volume_info", "vt_read_lbvh", "vt_instance_count", "vt_synchronize", "vt_last_error",
                     "vt_serve_init", "vt_serve_add_scene", "vt_serve_submit", "vt_serve_step", "vt_serve_read",
                     "vt_serve_status", "vt_serve_shutdown", "vt_enable_gbuffer", "vt_read_gbuffer",
                     "vt_dn_init", "vt_dn_layers", "vt_dn_param_count", "vt_dn_set_params", "vt_dn_get_params",
                     "vt_dn_grad", "vt_dn_features", "vt_dn_add_frame", "vt_dn_clear_frames", "vt_dn_train",
                     "vt_dn_denoise", "vt_dn_denoise_frame", "vt_dn_set_blend", "vt_dn_shutdown",
                     "vt_render_into")


class RenderTickInfo(C.Structure):
    """lib/common.h:89-92"""
    _fields_ = [("perspective", C.c_void_p), ("camera", C.c_void_p)]


class UserInput(C.Structure):
    """lib/common.h:145-151 == src/render.rs:37-51"""
    _fields_ = [("keys", C.c_uint8 * 6), ("mouse_x", C.c_double), ("mouse_y", C.c_double),
                ("last_mouse_x", C.c_double), ("last_mouse_y", C.c_double)]


class VtRenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("viewport_square", C.c_uint32),
                ("mode", C.c_uint32), ("spp", C.c_uint32), ("max_bounces", C.c_uint32), ("jitter", C.c_uint32),
                ("seed", C.c_uint32), ("frame", C.c_uint32), ("sun_dir", C.c_float * 3),
                ("sun_radiance", C.c_float * 3), ("sky", C.c_float * 3), ("stats", C.c_uint32)]


def declare(lib: C.CDLL) -> None:
    if getattr(lib, "_vt_declared", False):
        return
    lib.entry.restype = C.c_uint64
    lib.entry.argtypes = []
    lib.render_tick.restype = C.c_int32
    lib.render_tick.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(RenderTickInfo)]
    lib.get_input_data_pointer.restype = C.POINTER(UserInput)
    lib.get_input_data_pointer.argtypes = []
    lib.cleanup.restype = None
    lib.cleanup.argtypes = []
    lib.add_texture.restype = C.c_int32
    lib.add_texture.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.vt_add_textures.restype = C.c_int32
    lib.vt_add_textures.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.vt_lbvh_builds.restype = C.c_uint64
    lib.vt_lbvh_builds.argtypes = []
    lib.vt_refit_instances.restype = C.c_int32
    lib.vt_refit_instances.argtypes = [C.c_uint32]
    lib.vt_set_frame_split.restype = C.c_int32
    lib.vt_set_frame_split.argtypes = [C.c_uint32]
    lib.vt_set_persistent_blocks.restype = C.c_int32
    lib.vt_set_persistent_blocks.argtypes = [C.c_uint32]
    lib.vt_set_cursor.restype = None
    lib.vt_set_cursor.argtypes = [C.c_double, C.c_double]
    lib.start_update_instances.restype = C.POINTER(C.c_float)
    lib.start_update_instances.argtypes = [C.c_uint32]
    lib.end_update_instances.restype = C.c_int32
    lib.end_update_instances.argtypes = [C.c_uint32]
    lib.vt_set_params.restype = C.c_int32
    lib.vt_set_params.argtypes = [C.POINTER(VtRenderParams)]
    lib.vt_get_params.restype = C.c_int32
    lib.vt_get_params.argtypes = [C.POINTER(VtRenderParams)]
    lib.vt_add_volume_indexed.restype = C.c_int32
    lib.vt_add_volume_indexed.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.vt_render_into.restype = C.c_int32
    lib.vt_render_into.argtypes = [C.POINTER(RenderTickInfo), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
    lib.vt_enable_gbuffer.restype = C.c_int32
    lib.vt_enable_gbuffer.argtypes = [C.c_uint32]
    for name in ("vt_read_hits", "vt_read_radiance", "vt_read_bgra8", "vt_read_gbuffer"):
        f = getattr(lib, name)
        f.restype = C.c_int32
        f.argtypes = [C.c_void_p, C.c_uint64]
    lib.vt_read_stats.restype = C.c_int32
    lib.vt_read_stats.argtypes = [C.c_void_p]
    lib.vt_read_stage_stats.restype = C.c_int32
    lib.vt_read_stage_stats.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    lib.vt_bench_stages.restype = C.c_int32
    lib.vt_bench_stages.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_void_p,
                                    C.c_uint32, C.c_void_p]
    lib.vt_last_frame_ms.restype = C.c_float
    lib.vt_last_frame_ms.argtypes = []
    lib.vt_bench_frames.restype = C.c_int32
    lib.vt_bench_frames.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.vt_read_volume.restype = C.c_int32
    lib.vt_read_volume.argtypes = [C.c_int32, C.c_void_p, C.c_uint64]
    lib.vt_read_brick_map.restype = C.c_int32
    lib.vt_read_brick_map.argtypes = [C.c_int32, C.c_void_p, C.c_uint64]
    lib.vt_volume_info.restype = C.c_int32
    lib.vt_volume_info.argtypes = [C.c_int32, C.c_void_p]
    lib.vt_read_lbvh.restype = C.c_int32
    lib.vt_read_lbvh.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
    lib.vt_instance_count.restype = C.c_int32
    lib.vt_instance_count.argtypes = []
    lib.vt_synchronize.restype = C.c_int32
    lib.vt_synchronize.argtypes = []
    lib.vt_last_error.restype = C.c_char_p
    lib.vt_last_error.argtypes = []
    from . import denoise, serving
    serving.declare(lib)
    denoise.declare(lib)
    lib._vt_declared = True
