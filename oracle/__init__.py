"""THE ORACLE (test infrastructure): ctypes access to oracle/build/liboracle.so.

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
_lib = None
P = C.c_void_p


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        sig = {
            "or_pcg": (C.c_uint32, [C.c_uint32]),
            "or_pcg_uv": (None, [C.c_uint32, P]),
            "or_pack_gb": (C.c_uint64, [C.c_uint32] * 7),
            "or_unpack_gb": (None, [C.c_uint64, P]),
            "or_is_empty": (C.c_int, [C.c_uint64]),
            "or_quantize_depth": (C.c_uint32, [C.c_float] * 3),
            "or_dequantize_depth": (C.c_float, [C.c_uint32, C.c_float, C.c_float]),
            "or_oct_encode": (None, [P, P]),
            "or_oct_decode": (None, [C.c_uint32, C.c_uint32, P]),
            "or_control_points_for_lod": (C.c_uint32, [C.c_float]),
            "or_eval_cp": (None, [P, P, P, C.c_uint32, C.c_float, C.c_float, C.c_uint32, C.c_uint32, P]),
            "or_dom_optical_depth": (C.c_float, [P, C.c_float]),
            "or_clear": (None, [P, P, P, C.c_size_t]),
            "or_clear_dom": (None, [P, C.c_uint32]),
            "or_lod": (None, [P, P, P, P, P, C.c_int]),
            "or_scan": (C.c_uint32, [P, P, P, C.c_uint32, P]),
            "or_raster": (None, [P, P, P, P, P, C.c_uint32, P, P, P, P, P, C.c_float, P, C.c_int, C.c_int]),
            "or_shade": (None, [P, P, P, P, P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, P, C.c_float, P, P,
                                  C.c_int, C.c_int, C.c_uint32, C.c_uint32]),
            "or_range_falloff": (C.c_double, [C.c_double, C.c_double]),
            "or_scatter_tint": (None, [P, C.c_double, P]),
            "or_cube_face": (None, [P, P]),
            "or_cube_sample": (None, [P, C.c_uint32, C.c_uint32, P, C.c_double, P]),
            "or_cubemap_dir": (None, [P, P, P]),
            "or_probe_radiance": (None, [P, P, P, P, C.c_double, C.c_double, P]),
            "or_filter": (None, [P, P, P, P, P, P, P, C.c_int, C.c_int, C.c_uint32, C.c_uint32]),
            "or_azimuthal": (None, [C.c_double, C.c_double, P, P]),
            "or_dual_scatter": (None, [C.c_double, P, C.c_double, C.c_double, P]),
            "or_hair_bsdf": (None, [P, P, P, P, C.c_double, C.c_double, C.c_double, P]),
            "or_fresnel": (C.c_double, [C.c_double]),
            "or_scatter_sums": (None, [C.c_double, P, P]),
            "or_dual_terms": (None, [P, P, C.c_double, C.c_double, P]),
            "or_bsdf_parts": (None, [P, P, P, P, C.c_double, C.c_double, C.c_double, P]),
            "or_dom_visibility": (C.c_float, [P, P, P, C.c_float]),
            "or_spot_mask": (C.c_float, [P, P]),
            "or_distance_attenuation": (C.c_float, [C.c_float, C.c_float]),
            "or_reconstruct_world": (None, [P, C.c_int32, C.c_int32, C.c_float, P]),
            "or_project_world": (None, [P, P, P]),
            "or_screen_tangent": (None, [P, P, P, P]),
            "or_pixel_inputs": (None, [P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P]),
            "or_hair_visibility": (C.c_float, [P, P, C.c_uint32, P]),
            "or_composite": (None, [P, P, P, P, P]),
            "or_absorption_from_color": (None, [P, C.c_double, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype, fn.argtypes = res, args
        _lib = l
    return _lib


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return C.cast(C.byref(a), P)


class Frame:
    """CPU restatement of one groom's hair_pass over numpy buffers."""

    def __init__(self, mesh, width: int, height: int, dom_size: int):
        self.mesh = mesh
        self.width, self.height, self.dom_size = width, height, dom_size
        n = width * height
        self.center = np.empty(n, np.uint64)
        self.cons = np.empty(n, np.uint64)
        self.beta = np.empty(n, np.uint32)
        self.dom = np.zeros(dom_size * dom_size * 17, np.uint32)
        self.shaded = np.zeros((n, 4), np.float32)
        self.shaded_depth = np.zeros(n, np.float32)
        self.filtered = np.zeros((n, 4), np.float32)
        self.filtered_depth = np.zeros(n, np.float32)
        self.scene_depth = None
        self.scene_clear = 0.0

    def clear(self):
        lib().or_clear(_p(self.center), _p(self.cons), _p(self.beta), self.width * self.height)

    def clear_dom(self):
        lib().or_clear_dom(_p(self.dom), self.dom_size)

    def lod(self, params, control_points=0, strand_fraction=0.0, f64=False):
        from bevy_hair_amd._native import HmLodOverride
        out = np.zeros((max(1, self.mesh.bundle_count), 4), np.uint32)
        o = HmLodOverride(control_points, strand_fraction)
        lib().or_lod(_p(params), _p(self.mesh.bundles), _p(self.mesh.layers), _p(out), _p(o), int(f64))
        return out

    @staticmethod
    def scan(params, lod, capacity):
        refs = np.zeros((max(1, capacity), 4), np.uint32)
        ind = np.zeros(3, np.uint32)
        total = lib().or_scan(_p(params), _p(np.ascontiguousarray(lod)), _p(refs), capacity, _p(ind))
        return refs[:total].copy(), int(total)

    def raster(self, params, refs, f64=False, threads=1, write=True):
        counts = np.zeros(4, np.uint64)
        refs = np.ascontiguousarray(refs, dtype=np.uint32)
        if write:
            bufs = (self.center, self.cons, self.beta, self.dom)
        else:
            bufs = (None, None, None, self.dom if params.pass_mode[0] == 2 else None)
        lib().or_raster(_p(params), _p(self.mesh.bundles), _p(self.mesh.layers), _p(self.mesh.style), _p(refs),
                        len(refs), *(_p(b) for b in bufs), _p(self.scene_depth), self.scene_clear, _p(counts),
                        int(f64), threads)
        return counts

    # Rect lights (hm_rect_light array, count) and the environment probe
    # (hm_env_probe, its cube maps kept alive by the caller) the shading reads.
    rects = (None, 0)
    env = None

    def shade(self, params, points, n_points, dirs, n_dirs, exposure=1.0, f64=False, threads=1, rows=None):
        r0, r1 = rows if rows else (0, self.height)
        rects, n_rects = self.rects
        lib().or_shade(_p(params), _p(self.center), _p(self.beta), _p(self.dom), C.cast(points, P), n_points,
                       C.cast(dirs, P), n_dirs, C.cast(rects, P) if rects is not None else None, n_rects,
                       _p(self.env), exposure, _p(self.shaded), _p(self.shaded_depth), int(f64), threads, r0, r1)

    def composite(self, params, target_color, target_depth, filtered=None):
        """composite.wgsl: this frame's final colour (filtered if the camera
        params' filter flag is set, else shaded) into the view targets."""
        if filtered is None:
            filtered = bool(params.flags[1])
        color = self.filtered if filtered else self.shaded
        depth = self.filtered_depth if filtered else self.shaded_depth
        lib().or_composite(_p(params), _p(color), _p(depth), _p(target_color), _p(target_depth))

    def filter(self, params, f64=False, threads=1, rows=None):
        r0, r1 = rows if rows else (0, self.height)
        lib().or_filter(_p(params), _p(self.center), _p(self.cons), _p(self.shaded), _p(self.shaded_depth),
                        _p(self.filtered), _p(self.filtered_depth), int(f64), threads, r0, r1)

    def hair_pass(self, params, points, n_points, dirs, n_dirs, control_points=0, strand_fraction=0.0,
                  exposure=1.0, threads=1):
        """The per-groom sequence of pass.rs:379-619 (same order as hm_frame)."""
        cap = max(1, self.mesh.strand_count)
        cam = params[0]
        self.clear()
        lod = self.lod(cam, control_points, strand_fraction)
        refs, _ = self.scan(cam, lod, cap)
        self.raster(cam, refs, threads=threads)
        out = {"cam_lod": lod, "cam_refs": refs}
        if cam.flags[3]:
            self.clear_dom()
            dlod = self.lod(params[1])
            drefs, _ = self.scan(params[1], dlod, cap)
            self.raster(params[1], drefs, threads=threads)
            self.raster(params[2], drefs, threads=threads)
            out.update(dom_lod=dlod, dom_refs=drefs)
        self.shade(cam, points, n_points, dirs, n_dirs, exposure, threads=threads)
        if cam.flags[1]:
            self.filter(cam, threads=threads)
        return out
