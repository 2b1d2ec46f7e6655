"""ctypes binding of include/hair_mi.h (libhairmi.so, built in-tree).

There is no fallback: if the library is missing or fails to load, every
entry point raises. The structs mirror the header byte for byte.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhairmi.so")

u32 = C.c_uint32
f32 = C.c_float


class HmParams(C.Structure):
    """HairParams, 432 bytes (pass.rs:42-62)."""

    _fields_ = [
        ("clip_from_world", f32 * 16),
        ("world_from_clip", f32 * 16),
        ("world_from_model", f32 * 16),
        ("light_clip_from_world", f32 * 16),
        ("camera_pos", f32 * 4),
        ("camera_forward", f32 * 4),
        ("screen", f32 * 4),
        ("appearance", f32 * 4),
        ("filter_params", f32 * 4),
        ("albedo", f32 * 4),
        ("flags", u32 * 4),
        ("pass_mode", u32 * 4),
        ("light_eye", f32 * 4),
        ("light_forward", f32 * 4),
        ("dom_info", f32 * 4),
    ]


class HmPointLight(C.Structure):
    _fields_ = [
        ("position_radius", f32 * 4),
        ("color_inverse_square_range", f32 * 4),
        ("light_custom_data", f32 * 4),
        ("flags", u32),
        ("kind", u32),
        ("_pad", u32 * 2),
    ]


class HmDirectionalLight(C.Structure):
    _fields_ = [
        ("direction_to_light", f32 * 4),
        ("color", f32 * 4),
        ("flags", u32),
        ("_pad", u32 * 3),
    ]


class HmRectLight(C.Structure):
    _fields_ = [
        ("position", f32 * 4),
        ("right", f32 * 4),
        ("up", f32 * 4),
        ("color", f32 * 4),
        ("width", f32),
        ("height", f32),
        ("range", f32),
        ("_pad", f32),
    ]


class HmEnvProbe(C.Structure):
    _fields_ = [
        ("specular", C.c_void_p),
        ("specular_size", u32),
        ("specular_mips", u32),
        ("diffuse", C.c_void_p),
        ("diffuse_size", u32),
        ("view_rotation", f32 * 4),
        ("intensity", f32),
    ]


class HmGroomConfig(C.Structure):
    _fields_ = [
        ("world_from_model", f32 * 16),
        ("lambda_", f32),
        ("albedo_linear", f32 * 3),
        ("roughness", f32),
        ("tilt", f32),
        ("center_diameter", f32),
        ("lod", u32),
        ("filter", u32),
        ("ambient_occlusion", u32),
        ("deep_opacity", u32),
    ]


class HmView(C.Structure):
    _fields_ = [("clip_from_world", f32 * 16), ("world_from_view", f32 * 16)]


class HmGroomDesc(C.Structure):
    _fields_ = [
        ("bundles", C.c_void_p),
        ("bundle_count", u32),
        ("layers", C.c_void_p),
        ("layer_count", u32),
        ("style", C.c_void_p),
        ("style_count", u32),
        ("bounds_min", f32 * 3),
        ("bounds_max", f32 * 3),
    ]


class HmLodOverride(C.Structure):
    _fields_ = [("control_points", u32), ("strand_fraction", f32)]


class HmFrameOpts(C.Structure):
    _fields_ = [("exposure", f32), ("camera_lod", HmLodOverride), ("shading", u32), ("mlp_tolerance", f32)]


SHADING_PLAIN, SHADING_MLP = 0, 1
MLP_NW, MLP_NBIAS = 11264, 208
MLP_NOUT = 16  # network outputs (the dual-scattering terms)


class HmBatchView(C.Structure):
    _fields_ = [("groom", C.c_void_p), ("params", HmParams * 3), ("width", C.c_uint32), ("height", C.c_uint32),
                ("dom_size", C.c_uint32), ("opts", HmFrameOpts)]


class HmTuning(C.Structure):
    _fields_ = [("raster_flags", u32 * 3), ("raster_blocks_per_cu", u32 * 3), ("dom_tile", u32), ("cam_tile", u32),
                ("bin_box_limit", u32), ("bin_records", u32), ("bin_blocks_per_cu", u32),
                ("cam_chunk", u32), ("cam_bin_blocks_per_cu", u32), ("light_cull", u32),
                ("dom_busy_blocks_per_cu", u32), ("frame_serial", u32), ("batch_lanes", u32),
                ("filter_tiles", u32)]


RASTER_PRECHECK, RASTER_XCD = 1, 2


class HmShadowLight(C.Structure):
    _fields_ = [
        ("light_clip_from_world", f32 * 16),
        ("light_eye", f32 * 4),
        ("light_forward", f32 * 4),
        ("depth", f32 * 4),
    ]


assert C.sizeof(HmParams) == 432
assert C.sizeof(HmPointLight) == 64
assert C.sizeof(HmDirectionalLight) == 48
assert C.sizeof(HmShadowLight) == 112
assert C.sizeof(HmRectLight) == 80
assert C.sizeof(HmEnvProbe) == 48

# hm_buffer
BUF_CENTER, BUF_CONSERVATIVE, BUF_BETA, BUF_SHADED, BUF_SHADED_DEPTH = 0, 1, 2, 3, 4
BUF_FILTERED, BUF_FILTERED_DEPTH, BUF_DOM, BUF_SCENE_DEPTH = 5, 6, 7, 8
BUF_CAMERA_LOD, BUF_CAMERA_REFS, BUF_CAMERA_INDIRECT = 16, 17, 18
BUF_DOM_LOD, BUF_DOM_REFS, BUF_DOM_INDIRECT = 19, 20, 21
BUF_CAM_BIN_STATUS, BUF_DOM_BIN_STATUS = 24, 25
# Status words of a binned pass (hm_kernels.h): 16 header words, then the
# shade and filter sub-list counters of the camera tiles.
LIST_SUBS = 64
BIN_SHADE_SUBS = 16
BIN_META_WORDS = BIN_SHADE_SUBS + 2 * LIST_SUBS
PASS_CAMERA, PASS_DOM = 0, 1
KERNELS = ["clear", "cam_lod", "cam_scan", "cam_raster", "clear_dom", "dom_lod", "dom_scan",
           "dom_depth", "dom_opacity", "shade", "filter", "dom_bin", "dom_tiles", "cam_bin", "cam_tiles",
           "cam_finish", "dom_long_depth", "dom_long_opacity"]

# The exported C ABI: name -> (restype, argtypes). Every symbol include/hair_mi.h declares.
P = C.c_void_p
ABI = {
    "hm_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "hm_destroy": (None, [P]),
    "hm_status_string": (C.c_char_p, [C.c_int]),
    "hm_last_error": (C.c_char_p, [P]),
    "hm_groom_upload": (C.c_int, [P, C.POINTER(HmGroomDesc), C.POINTER(P)]),
    "hm_groom_free": (None, [P, P]),
    "hm_groom_bundle_count": (u32, [P]),
    "hm_groom_strand_count": (u32, [P]),
    "hm_targets": (C.c_int, [P, u32, u32, u32]),
    "hm_set_scene_depth": (C.c_int, [P, P, f32]),
    "hm_set_lights": (C.c_int, [P, P, u32, P, u32]),
    "hm_set_rect_lights": (C.c_int, [P, P, u32]),
    "hm_set_env_probe": (C.c_int, [P, P]),
    "hm_make_params": (C.c_int, [C.POINTER(HmView), C.POINTER(HmGroomConfig), f32 * 3, f32 * 3, u32,
                                 f32 * 3, u32, u32, u32, C.POINTER(HmParams)]),
    "hm_key_direction": (None, [P, u32, f32 * 3]),
    "hm_shadow_light_from_params": (None, [C.POINTER(HmParams), C.POINTER(HmShadowLight)]),
    "hm_clear": (C.c_int, [P]),
    "hm_clear_dom": (C.c_int, [P]),
    "hm_lod": (C.c_int, [P, P, C.POINTER(HmParams), C.c_int, C.POINTER(HmLodOverride)]),
    "hm_scan": (C.c_int, [P, P, C.POINTER(HmParams), C.c_int]),
    "hm_raster": (C.c_int, [P, P, C.POINTER(HmParams), C.c_int]),
    "hm_shade": (C.c_int, [P, C.POINTER(HmParams), f32]),
    "hm_filter": (C.c_int, [P, C.POINTER(HmParams)]),
    "hm_gbuffer": (C.c_int, [P, P, C.POINTER(HmParams)]),
    "hm_dom": (C.c_int, [P, P, C.POINTER(HmParams), C.POINTER(HmParams)]),
    "hm_raster_count": (C.c_int, [P, P, C.POINTER(HmParams), C.c_int, C.c_uint64 * 4]),
    "hm_frame": (C.c_int, [P, P, C.POINTER(HmParams), C.POINTER(HmFrameOpts)]),
    "hm_read": (C.c_int, [P, P, C.c_int, P, C.c_size_t]),
    "hm_device_ptr": (C.c_int, [P, P, C.c_int, C.POINTER(P), C.POINTER(C.c_size_t)]),
    "hm_sync": (C.c_int, [P]),
    "hm_timing_enable": (C.c_int, [P, C.c_int]),
    "hm_timing_read": (C.c_int, [P, C.POINTER(f32), u32]),
    "hm_time_frames": (C.c_int, [P, P, C.POINTER(HmParams), C.POINTER(HmFrameOpts), u32, C.POINTER(f32)]),
    "hm_get_tuning": (C.c_int, [P, C.POINTER(HmTuning)]),
    "hm_set_tuning": (C.c_int, [P, C.POINTER(HmTuning)]),
    "hm_autotune": (C.c_int, [P, P, C.POINTER(HmParams), C.POINTER(HmFrameOpts), u32, C.POINTER(HmTuning)]),
    "hm_bench_copy": (C.c_int, [P, C.c_size_t, u32, C.POINTER(f32)]),
    "hm_bench_atomics": (C.c_int, [P, C.c_size_t, u32, C.c_uint64, C.POINTER(f32)]),
    "hm_bench_lds_atomics": (C.c_int, [P, u32, u32, C.c_uint64, C.POINTER(f32)]),
    "hm_mlp_init": (C.c_int, [P, u32]),
    "hm_mlp_set_weights": (C.c_int, [P, P, P]),
    "hm_mlp_get_weights": (C.c_int, [P, P, P]),
    "hm_mlp_forward": (C.c_int, [P, P, u32, P]),
    "hm_mlp_gradients": (C.c_int, [P, P, P, u32, P, C.POINTER(f32), P, P]),
    "hm_mlp_reserve_samples": (C.c_int, [P, u32]),
    "hm_mlp_collect": (C.c_int, [P, C.POINTER(HmParams), f32, u32, C.POINTER(u32)]),
    "hm_mlp_read_samples": (C.c_int, [P, P, P, u32]),
    "hm_mlp_set_samples": (C.c_int, [P, P, P, u32]),
    "hm_mlp_set_range": (C.c_int, [P, P, P]),
    "hm_mlp_train": (C.c_int, [P, u32, u32, f32, P, u32, C.POINTER(f32)]),
    "hm_shade_mlp": (C.c_int, [P, C.POINTER(HmParams), f32, f32, C.POINTER(u32)]),
    "hm_batch_create": (C.c_int, [P, C.POINTER(HmBatchView), u32, C.POINTER(C.c_void_p)]),
    "hm_batch_run": (C.c_int, [P, P, u32]),
    "hm_batch_time": (C.c_int, [P, P, u32, C.POINTER(f32)]),
    "hm_batch_read": (C.c_int, [P, P, u32, P, P]),
    "hm_batch_free": (None, [P, P]),
    "hm_scan_noalloc": (C.c_int, [P, P, u32, P, u32, P]),
    "hm_bind_scene_depth": (C.c_int, [P, P, f32]),
    "hm_composite": (C.c_int, [P, C.POINTER(HmParams), P, P]),
    "hm_shadow_factors": (C.c_int, [P, C.POINTER(HmShadowLight), P, u32, P, u32, P]),
    "hm_bake": (C.c_int, [P, u32, P, u32, C.POINTER(C.c_void_p)]),
    "hm_bake_get": (C.c_int, [P, C.POINTER(HmGroomDesc)]),
    "hm_bake_free": (None, [P]),
}
SCAN_NOALLOC_MAX_BUNDLES = 65536

_lib = None


class NativeError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Loads libhairmi.so once; raises NativeError if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} is not built (run bevy_hair_amd.build or __graft_entry__.build())")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int, ctx=None, what: str = "") -> None:
    if status != 0:
        l = lib()
        msg = l.hm_status_string(status).decode()
        detail = l.hm_last_error(ctx).decode() if ctx else ""
        raise NativeError(f"{what}: {msg} ({status}) {detail}")


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def params_array(params) -> C.Array:
    arr = (HmParams * 3)()
    for i in range(3):
        arr[i] = params[i]
    return arr
