"""Host-side mirror of bevy_hair's plugin surface over the C ABI.

HairGroom           src/lib.rs:32-64 (same fields and defaults)
Renderer.upload     GpuGroom::upload, src/render/pass.rs:98-167
Renderer.hair_pass  hair_pass's per-groom sequence, src/render/pass.rs:375-643
Renderer.shadow     HairShadow publishing of the DOM + ShadowLight, pass.rs:571-580

Everything runs on the GPU through libhairmi.so; there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _native as N
from .groom import HairMesh
from .scene import View, cols


def srgb_to_linear(c: float) -> float:
    """Bevy Color::srgb(..).to_linear() per channel."""
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


@dataclasses.dataclass
class HairGroom:
    """A groom instance (lib.rs:45-64): shading parameters and feature toggles."""

    lambda_: float = 3.0
    albedo_srgb: tuple = (0.16, 0.07, 0.035)
    albedo_linear: tuple | None = None  # overrides albedo_srgb when set
    roughness: float = 0.2
    tilt: float = 0.08
    lod: bool = True
    filter: bool = True
    ambient_occlusion: bool = True
    deep_opacity: bool = True
    center_diameter: float = 1.0
    world_from_model: np.ndarray = dataclasses.field(default_factory=lambda: np.eye(4))

    def config(self) -> N.HmGroomConfig:
        g = N.HmGroomConfig()
        g.world_from_model[:] = cols(self.world_from_model)
        g.lambda_ = self.lambda_
        lin = self.albedo_linear or tuple(srgb_to_linear(c) for c in self.albedo_srgb)
        g.albedo_linear[:] = [float(x) for x in lin]
        g.roughness, g.tilt, g.center_diameter = self.roughness, self.tilt, self.center_diameter
        g.lod, g.filter = int(self.lod), int(self.filter)
        g.ambient_occlusion, g.deep_opacity = int(self.ambient_occlusion), int(self.deep_opacity)
        return g


def point_light(d: dict) -> N.HmPointLight:
    L = N.HmPointLight()
    L.position_radius[:] = d["position_radius"]
    L.color_inverse_square_range[:] = d["color_inverse_square_range"]
    L.light_custom_data[:] = d.get("light_custom_data", [0, 0, 0, 0])
    L.flags, L.kind = d.get("flags", 0), d.get("kind", 0)
    return L


def directional_light(d: dict) -> N.HmDirectionalLight:
    L = N.HmDirectionalLight()
    L.direction_to_light[:] = d["direction_to_light"]
    L.color[:] = d["color"]
    L.flags = d.get("flags", 0)
    return L


def rect_light(d: dict) -> N.HmRectLight:
    L = N.HmRectLight()
    L.position[:] = d["position"]
    L.right[:] = d["right"]
    L.up[:] = d["up"]
    L.color[:] = d["color"]
    L.width, L.height, L.range = d["width"], d["height"], d["range"]
    return L


def rect_arrays(rects: list[dict]):
    return (N.HmRectLight * max(1, len(rects)))(*[rect_light(r) for r in rects]), len(rects)


class EnvProbeDesc:
    """hm_env_probe over an EnvironmentMapLight, holding the packed maps."""

    def __init__(self, env):
        env.validate()
        self.spec, self.diff = env.packed()
        d = N.HmEnvProbe()
        d.specular, d.specular_size, d.specular_mips = N.ptr(self.spec), env.specular[0].shape[1], len(env.specular)
        d.diffuse, d.diffuse_size = N.ptr(self.diff), env.diffuse.shape[1]
        d.view_rotation[:] = [float(c) for c in env.rotation]
        d.intensity = float(env.intensity)
        self.desc = d


def light_arrays(points: list[dict], dirs: list[dict]):
    pa = (N.HmPointLight * max(1, len(points)))(*[point_light(p) for p in points])
    da = (N.HmDirectionalLight * max(1, len(dirs)))(*[directional_light(d) for d in dirs])
    return pa, len(points), da, len(dirs)


class GpuGroom:
    """Device copy of one HairMesh (owned by a Renderer)."""

    def __init__(self, renderer: "Renderer", mesh: HairMesh):
        mesh.validate()
        self.renderer, self.mesh = renderer, mesh
        self._arrays = (np.ascontiguousarray(mesh.bundles), np.ascontiguousarray(mesh.layers),
                        np.ascontiguousarray(mesh.style))
        d = N.HmGroomDesc()
        d.bundles, d.bundle_count = N.ptr(self._arrays[0]), mesh.bundle_count
        d.layers, d.layer_count = N.ptr(self._arrays[1]), len(mesh.layers)
        d.style, d.style_count = N.ptr(self._arrays[2]), len(mesh.style)
        d.bounds_min[:] = mesh.bounds_min.tolist()
        d.bounds_max[:] = mesh.bounds_max.tolist()
        h = C.c_void_p()
        N.check(N.lib().hm_groom_upload(renderer.ctx, C.byref(d), C.byref(h)), renderer.ctx, "hm_groom_upload")
        self.handle = h

    @property
    def bundle_count(self) -> int:
        return int(N.lib().hm_groom_bundle_count(self.handle))

    def free(self):
        if self.handle:
            N.lib().hm_groom_free(self.renderer.ctx, self.handle)
            self.handle = None


class Batch:
    """A captured multi-view frame owned by a Renderer (hm_batch_*)."""

    def __init__(self, renderer: "Renderer", handle, sizes):
        self.renderer, self.handle, self.sizes = renderer, handle, sizes

    def run(self, reps: int = 1):
        self.renderer._check(N.lib().hm_batch_run(self.renderer.ctx, self.handle, reps), "hm_batch_run")

    def time(self, reps: int) -> float:
        ms = C.c_float()
        self.renderer._check(N.lib().hm_batch_time(self.renderer.ctx, self.handle, reps, C.byref(ms)), "hm_batch_time")
        return float(ms.value)

    def read(self, view: int):
        w, h = self.sizes[view]
        rgba = np.empty((w * h, 4), np.float32)
        depth = np.empty(w * h, np.float32)
        self.renderer._check(N.lib().hm_batch_read(self.renderer.ctx, self.handle, view, N.ptr(rgba), N.ptr(depth)),
                             "hm_batch_read")
        return rgba, depth

    def free(self):
        if self.handle:
            N.lib().hm_batch_free(self.renderer.ctx, self.handle)
            self.handle = None


class Renderer:
    """One device context (hm_ctx): targets, lights, grooms, frames."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = N.lib().hm_create(device, C.byref(h))
        N.check(st, None, f"hm_create(device={device})")
        self.ctx = h
        self.width = self.height = self.dom_size = 0

    def close(self):
        if self.ctx:
            N.lib().hm_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st, what):
        N.check(st, self.ctx, what)

    def upload(self, mesh: HairMesh) -> GpuGroom:
        return GpuGroom(self, mesh)

    def targets(self, width: int, height: int, dom_size: int):
        self._check(N.lib().hm_targets(self.ctx, width, height, dom_size), "hm_targets")
        self.width, self.height, self.dom_size = width, height, dom_size

    def set_scene_depth(self, depth: np.ndarray | None, clear_value: float = 0.0):
        if depth is None:
            self._check(N.lib().hm_set_scene_depth(self.ctx, None, clear_value), "hm_set_scene_depth")
        else:
            d = np.ascontiguousarray(depth, dtype=np.float32)
            assert d.size == self.width * self.height
            self._check(N.lib().hm_set_scene_depth(self.ctx, N.ptr(d), clear_value), "hm_set_scene_depth")

    def bind_scene_depth(self, device_ptr: int | None, clear_value: float = 0.0):
        """The scene depth as a caller-owned device buffer (no copy); None unbinds."""
        self._check(N.lib().hm_bind_scene_depth(self.ctx, device_ptr, clear_value), "hm_bind_scene_depth")

    def composite(self, params, color_ptr: int, depth_ptr: int):
        """composite.wgsl into device view targets float4[W*H] / float[W*H]."""
        self._check(N.lib().hm_composite(self.ctx, C.byref(params), color_ptr, depth_ptr), "hm_composite")

    def shadow_factors(self, light: N.HmShadowLight, dom_ptr: int, dom_size: int, points_ptr: int, count: int,
                       out_ptr: int):
        """HairShadowMaterial's DOM lookup for `count` float4 world points (device buffers)."""
        self._check(N.lib().hm_shadow_factors(self.ctx, C.byref(light), dom_ptr, dom_size, points_ptr, count, out_ptr),
                    "hm_shadow_factors")

    def set_lights(self, points: list[dict], directionals: list[dict]):
        pa, np_, da, nd = light_arrays(points, directionals)
        self._check(N.lib().hm_set_lights(self.ctx, pa, np_, da, nd), "hm_set_lights")
        self.directionals = directionals

    def set_rect_lights(self, rects: list[dict]):
        """The view's rect lights (shade.wesl:503-518); [] removes them."""
        ra, n = rect_arrays(rects)
        self._check(N.lib().hm_set_rect_lights(self.ctx, ra, n), "hm_set_rect_lights")

    def set_environment(self, env):
        """The view's EnvironmentMapLight (probe_radiance x AO is added to
        every hit pixel); None removes it."""
        if env is None:
            self._check(N.lib().hm_set_env_probe(self.ctx, None), "hm_set_env_probe")
            return
        d = EnvProbeDesc(env)
        self._check(N.lib().hm_set_env_probe(self.ctx, C.byref(d.desc)), "hm_set_env_probe")

    @staticmethod
    def make_params(view: View, groom: HairGroom, mesh: HairMesh, dom_size: int, directionals: list[dict]):
        """The camera / DOM-depth / DOM-opacity HairParams (pass.rs:399-453)."""
        v = N.HmView()
        v.clip_from_world[:] = cols(view.clip_from_world)
        v.world_from_view[:] = cols(view.world_from_view)
        cfg = groom.config()
        da = (N.HmDirectionalLight * max(1, len(directionals)))(*[directional_light(d) for d in directionals])
        key = (C.c_float * 3)()
        N.lib().hm_key_direction(da, len(directionals), key)
        out = (N.HmParams * 3)()
        bmin = (C.c_float * 3)(*mesh.bounds_min.tolist())
        bmax = (C.c_float * 3)(*mesh.bounds_max.tolist())
        st = N.lib().hm_make_params(C.byref(v), C.byref(cfg), bmin, bmax, max(1, mesh.bundle_count), key,
                                    view.width, view.height, dom_size, out)
        N.check(st, None, "hm_make_params")
        return out

    # -- kernel-level passes (one per WGSL dispatch) --
    def clear(self):
        self._check(N.lib().hm_clear(self.ctx), "hm_clear")

    def clear_dom(self):
        self._check(N.lib().hm_clear_dom(self.ctx), "hm_clear_dom")

    def lod(self, g: GpuGroom, params, pass_: int, control_points: int = 0, strand_fraction: float = 0.0):
        o = N.HmLodOverride(control_points, strand_fraction)
        self._check(N.lib().hm_lod(self.ctx, g.handle, C.byref(params), pass_, C.byref(o)), "hm_lod")

    def scan(self, g: GpuGroom, params, pass_: int):
        self._check(N.lib().hm_scan(self.ctx, g.handle, C.byref(params), pass_), "hm_scan")

    def scan_noalloc(self, lod_ptr: int, bundle_count: int, refs_ptr: int, capacity: int, indirect_ptr: int):
        """hm_scan over caller-owned device buffers, allocating nothing."""
        self._check(N.lib().hm_scan_noalloc(self.ctx, lod_ptr, bundle_count, refs_ptr, capacity, indirect_ptr),
                    "hm_scan_noalloc")

    def device_ptr(self, which: int, groom: GpuGroom | None = None) -> tuple[int, int]:
        """Device address and size of a library buffer (hm_device_ptr)."""
        p, n = C.c_void_p(), C.c_size_t()
        self._check(N.lib().hm_device_ptr(self.ctx, groom.handle if groom else None, which, C.byref(p),
                                          C.byref(n)), "hm_device_ptr")
        return int(p.value or 0), int(n.value)

    def raster(self, g: GpuGroom, params, pass_: int):
        self._check(N.lib().hm_raster(self.ctx, g.handle, C.byref(params), pass_), "hm_raster")

    def gbuffer_pass(self, g: GpuGroom, params):
        """clear + camera raster in one call (binned when tuned so)."""
        self._check(N.lib().hm_gbuffer(self.ctx, g.handle, C.byref(params)), "hm_gbuffer")

    def dom(self, g: GpuGroom, depth_params, opacity_params):
        """clear_dom + DOM depth + DOM opacity in one call (binned when tuned so)."""
        self._check(N.lib().hm_dom(self.ctx, g.handle, C.byref(depth_params), C.byref(opacity_params)), "hm_dom")

    def bin_status(self, pass_: int = 0) -> dict:
        """Status words of the last binned hm_gbuffer (pass 0) or hm_dom (pass 1)."""
        s = self.read(N.BUF_CAM_BIN_STATUS if pass_ == 0 else N.BUF_DOM_BIN_STATUS, np.uint32, N.BIN_META_WORDS)
        out = {"records": int(s[0]), "binned": bool(s[1]), "fallback": bool(s[2]), "long": bool(s[3]),
               "long_count": int(s[4])}
        if pass_ == 0:
            # words 16-79 / 80-143: the shade / filter sub-list counters the camera
            # tiles of a frame fill (zero after a pass that emits no lists); the
            # DOM pass keeps its tile workgroups' cursors there
            subs = N.BIN_SHADE_SUBS
            out["shade_pixels"] = int(s[subs:subs + N.LIST_SUBS].astype(np.int64).sum())
            out["filter_pixels"] = int(s[subs + N.LIST_SUBS:subs + 2 * N.LIST_SUBS].astype(np.int64).sum())
        return out

    def dom_bin_status(self) -> dict:
        return self.bin_status(1)

    def raster_count(self, g: GpuGroom, params, pass_: int) -> np.ndarray:
        c = (C.c_uint64 * 4)()
        self._check(N.lib().hm_raster_count(self.ctx, g.handle, C.byref(params), pass_, c), "hm_raster_count")
        return np.array(list(c), dtype=np.uint64)

    def shade(self, params, exposure: float = 1.0):
        self._check(N.lib().hm_shade(self.ctx, C.byref(params), exposure), "hm_shade")

    def filter(self, params):
        self._check(N.lib().hm_filter(self.ctx, C.byref(params)), "hm_filter")

    # -- frame driver --
    @staticmethod
    def frame_opts(exposure=1.0, control_points=0, strand_fraction=0.0, mlp: bool = False,
                   mlp_tolerance: float = 0.05) -> N.HmFrameOpts:
        return N.HmFrameOpts(exposure, N.HmLodOverride(control_points, strand_fraction),
                             N.SHADING_MLP if mlp else N.SHADING_PLAIN, mlp_tolerance)

    def hair_pass(self, g: GpuGroom, params, opts: N.HmFrameOpts | None = None):
        opts = opts or self.frame_opts()
        self._check(N.lib().hm_frame(self.ctx, g.handle, params, C.byref(opts)), "hm_frame")

    def batch(self, views) -> "Batch":
        """A captured multi-view frame (hm_batch_create): `views` are
        (GpuGroom, params[3], width, height, dom_size, HmFrameOpts) tuples."""
        arr = (N.HmBatchView * len(views))()
        for i, (g, params, w, h, dom, opts) in enumerate(views):
            arr[i].groom = g.handle
            for k in range(3):
                arr[i].params[k] = params[k]
            arr[i].width, arr[i].height, arr[i].dom_size = w, h, dom
            arr[i].opts = opts
        out = C.c_void_p()
        self._check(N.lib().hm_batch_create(self.ctx, arr, len(views), C.byref(out)), "hm_batch_create")
        return Batch(self, out.value, [(w, h) for (_, _, w, h, _, _) in views])

    def time_frames(self, g: GpuGroom, params, opts, frames: int) -> float:
        ms = C.c_float()
        self._check(N.lib().hm_time_frames(self.ctx, g.handle, params, C.byref(opts), frames, C.byref(ms)),
                    "hm_time_frames")
        return float(ms.value)

    def timing(self, on: bool):
        self._check(N.lib().hm_timing_enable(self.ctx, int(on)), "hm_timing_enable")

    def kernel_times(self) -> dict:
        ms = (C.c_float * len(N.KERNELS))()
        self._check(N.lib().hm_timing_read(self.ctx, ms, len(N.KERNELS)), "hm_timing_read")
        return dict(zip(N.KERNELS, list(ms)))

    def sync(self):
        self._check(N.lib().hm_sync(self.ctx), "hm_sync")

    # -- readback --
    def read(self, which: int, dtype, count: int, groom: GpuGroom | None = None) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        self._check(N.lib().hm_read(self.ctx, groom.handle if groom else None, which, N.ptr(out), out.nbytes),
                    "hm_read")
        return out

    def gbuffer(self):
        n = self.width * self.height
        return (self.read(N.BUF_CENTER, np.uint64, n), self.read(N.BUF_CONSERVATIVE, np.uint64, n),
                self.read(N.BUF_BETA, np.uint32, n))

    def dom_map(self) -> np.ndarray:
        return self.read(N.BUF_DOM, np.uint32, self.dom_size * self.dom_size * 17)

    def shaded(self):
        n = self.width * self.height
        return (self.read(N.BUF_SHADED, np.float32, n * 4).reshape(n, 4),
                self.read(N.BUF_SHADED_DEPTH, np.float32, n))

    def filtered(self):
        n = self.width * self.height
        return (self.read(N.BUF_FILTERED, np.float32, n * 4).reshape(n, 4),
                self.read(N.BUF_FILTERED_DEPTH, np.float32, n))

    def lod_records(self, g: GpuGroom, pass_: int) -> np.ndarray:
        which = N.BUF_CAMERA_LOD if pass_ == 0 else N.BUF_DOM_LOD
        return self.read(which, np.uint32, g.bundle_count * 4, g).reshape(-1, 4)

    def refs(self, g: GpuGroom, pass_: int) -> tuple[np.ndarray, int]:
        ind = self.read(N.BUF_CAMERA_INDIRECT if pass_ == 0 else N.BUF_DOM_INDIRECT, np.uint32, 4, g)
        total = int(ind[0])
        cap = max(1, g.mesh.strand_count)
        refs = self.read(N.BUF_CAMERA_REFS if pass_ == 0 else N.BUF_DOM_REFS, np.uint32, cap * 4, g).reshape(-1, 4)
        return refs[:total], total

    # -- tuning --
    @staticmethod
    def _tuning_dict(t: N.HmTuning) -> dict:
        return {"raster_flags": list(t.raster_flags), "raster_blocks_per_cu": list(t.raster_blocks_per_cu),
                "dom_tile": int(t.dom_tile), "cam_tile": int(t.cam_tile), "bin_box_limit": int(t.bin_box_limit),
                "bin_records": int(t.bin_records), "bin_blocks_per_cu": int(t.bin_blocks_per_cu),
                "cam_chunk": int(t.cam_chunk), "cam_bin_blocks_per_cu": int(t.cam_bin_blocks_per_cu),
                "light_cull": int(t.light_cull), "dom_busy_blocks_per_cu": int(t.dom_busy_blocks_per_cu),
                "frame_serial": int(t.frame_serial), "batch_lanes": int(t.batch_lanes),
                "filter_tiles": int(t.filter_tiles)}

    def get_tuning(self) -> dict:
        t = N.HmTuning()
        self._check(N.lib().hm_get_tuning(self.ctx, C.byref(t)), "hm_get_tuning")
        return self._tuning_dict(t)

    def set_tuning(self, raster_flags=None, raster_blocks_per_cu=None, dom_tile=None, bin_box_limit=None,
                   bin_records=None, bin_blocks_per_cu=None, cam_tile=None, cam_chunk=None,
                   cam_bin_blocks_per_cu=None, light_cull=None, dom_busy_blocks_per_cu=None, frame_serial=None,
                   batch_lanes=None, filter_tiles=None):
        """Changes the given launch choices; None keeps the current one."""
        cur = self.get_tuning()
        if raster_flags is not None:
            cur["raster_flags"] = [raster_flags] * 3 if isinstance(raster_flags, int) else list(raster_flags)
        if raster_blocks_per_cu is not None:
            b = raster_blocks_per_cu
            cur["raster_blocks_per_cu"] = [b] * 3 if isinstance(b, int) else list(b)
        for k, v in (("dom_tile", dom_tile), ("cam_tile", cam_tile), ("bin_box_limit", bin_box_limit),
                     ("bin_records", bin_records),
                     ("bin_blocks_per_cu", bin_blocks_per_cu), ("cam_chunk", cam_chunk),
                     ("cam_bin_blocks_per_cu", cam_bin_blocks_per_cu), ("light_cull", light_cull),
                     ("dom_busy_blocks_per_cu", dom_busy_blocks_per_cu), ("frame_serial", frame_serial),
                     ("batch_lanes", batch_lanes), ("filter_tiles", filter_tiles)):
            if v is not None:
                cur[k] = int(v)
        t = N.HmTuning((C.c_uint32 * 3)(*cur["raster_flags"]), (C.c_uint32 * 3)(*cur["raster_blocks_per_cu"]),
                       cur["dom_tile"], cur["cam_tile"], cur["bin_box_limit"], cur["bin_records"],
                       cur["bin_blocks_per_cu"], cur["cam_chunk"], cur["cam_bin_blocks_per_cu"], cur["light_cull"],
                       cur["dom_busy_blocks_per_cu"], cur["frame_serial"], cur["batch_lanes"],
                       cur["filter_tiles"])
        self._check(N.lib().hm_set_tuning(self.ctx, C.byref(t)), "hm_set_tuning")

    def autotune(self, g: GpuGroom, params, opts, reps: int = 3) -> dict:
        t = N.HmTuning()
        self._check(N.lib().hm_autotune(self.ctx, g.handle, params, C.byref(opts), reps, C.byref(t)), "hm_autotune")
        return self._tuning_dict(t)

    # -- roofline microbenchmarks --
    def bench_copy(self, nbytes: int, reps: int = 5) -> float:
        v = C.c_float()
        self._check(N.lib().hm_bench_copy(self.ctx, nbytes, reps, C.byref(v)), "hm_bench_copy")
        return float(v.value)

    def bench_atomics(self, target_bytes: int, kind: int, count: int) -> float:
        v = C.c_float()
        self._check(N.lib().hm_bench_atomics(self.ctx, target_bytes, kind, count, C.byref(v)), "hm_bench_atomics")
        return float(v.value)

    def bench_lds_atomics(self, lds_bytes: int, kind: int, count: int) -> float:
        v = C.c_float()
        self._check(N.lib().hm_bench_lds_atomics(self.ctx, lds_bytes, kind, count, C.byref(v)),
                    "hm_bench_lds_atomics")
        return float(v.value)

    # -- learned shading --
    def mlp_init(self, seed: int = 0):
        self._check(N.lib().hm_mlp_init(self.ctx, seed), "hm_mlp_init")

    def mlp_set_weights(self, w: np.ndarray, b: np.ndarray):
        w = np.ascontiguousarray(w, np.float32).reshape(-1)
        b = np.ascontiguousarray(b, np.float32).reshape(-1)
        assert w.size == N.MLP_NW and b.size == N.MLP_NBIAS
        self._check(N.lib().hm_mlp_set_weights(self.ctx, N.ptr(w), N.ptr(b)), "hm_mlp_set_weights")

    def mlp_get_weights(self):
        w = np.empty(N.MLP_NW, np.float32)
        b = np.empty(N.MLP_NBIAS, np.float32)
        self._check(N.lib().hm_mlp_get_weights(self.ctx, N.ptr(w), N.ptr(b)), "hm_mlp_get_weights")
        return w, b

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], 16), np.float32)
        self._check(N.lib().hm_mlp_forward(self.ctx, N.ptr(x), x.shape[0], N.ptr(y)), "hm_mlp_forward")
        return y

    def mlp_gradients(self, x: np.ndarray, t: np.ndarray, scale: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        scale = np.ascontiguousarray(scale, np.float32)
        dw = np.empty(N.MLP_NW, np.float32)
        db = np.empty(N.MLP_NBIAS, np.float32)
        loss = C.c_float()
        self._check(N.lib().hm_mlp_gradients(self.ctx, N.ptr(x), N.ptr(t), x.shape[0], N.ptr(scale), C.byref(loss),
                                              N.ptr(dw), N.ptr(db)), "hm_mlp_gradients")
        return float(loss.value), dw, db

    def mlp_reserve_samples(self, capacity: int):
        self._check(N.lib().hm_mlp_reserve_samples(self.ctx, capacity), "hm_mlp_reserve_samples")

    def mlp_collect(self, params, keep: float, seed: int) -> int:
        total = C.c_uint32()
        self._check(N.lib().hm_mlp_collect(self.ctx, C.byref(params), keep, seed, C.byref(total)), "hm_mlp_collect")
        return int(total.value)

    def mlp_read_samples(self, n: int):
        x = np.empty((n, 32), np.float32)
        t = np.empty((n, N.MLP_NOUT), np.float32)
        self._check(N.lib().hm_mlp_read_samples(self.ctx, N.ptr(x), N.ptr(t), n), "hm_mlp_read_samples")
        return x, t

    def mlp_set_samples(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        assert x.shape[1] == 32 and t.shape[1] == N.MLP_NOUT and x.shape[0] == t.shape[0]
        self._check(N.lib().hm_mlp_set_samples(self.ctx, N.ptr(x), N.ptr(t), x.shape[0]), "hm_mlp_set_samples")

    def mlp_set_range(self, lo, hi):
        if lo is None:
            self._check(N.lib().hm_mlp_set_range(self.ctx, None, None), "hm_mlp_set_range")
            return
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        self._check(N.lib().hm_mlp_set_range(self.ctx, N.ptr(lo), N.ptr(hi)), "hm_mlp_set_range")

    def mlp_train(self, epochs: int, batch: int, lr: float, scale, seed: int = 0) -> float:
        scale = np.ascontiguousarray(scale, np.float32)
        loss = C.c_float()
        self._check(N.lib().hm_mlp_train(self.ctx, epochs, batch, lr, N.ptr(scale), seed, C.byref(loss)),
                    "hm_mlp_train")
        return float(loss.value)

    def shade_mlp(self, params, exposure: float = 1.0, tolerance: float = 0.05, count: bool = True):
        n = C.c_uint32()
        self._check(N.lib().hm_shade_mlp(self.ctx, C.byref(params), exposure, tolerance,
                                         C.byref(n) if count else None), "hm_shade_mlp")
        return int(n.value) if count else None

    def shadow_light(self, dom_params) -> N.HmShadowLight:
        out = N.HmShadowLight()
        N.lib().hm_shadow_light_from_params(C.byref(dom_params), C.byref(out))
        return out
