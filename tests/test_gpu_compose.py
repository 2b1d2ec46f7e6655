"""After the frame: the shadow receiver lookup and the composite, HIP path
against the oracle.

- hm_shadow_factors (hair_shadow.wesl:30-75) for world points in and around
  the groom and on a ground plane below it, maps of 512^2 and 2048^2,
  within 1e-3 of the oracle's visibility;
- hm_composite (composite.wgsl + pipeline.rs:299-339) of two instances of
  one groom into shared view targets, the second groom's raster reading the
  view depth the first composited (hm_bind_scene_depth: no copy), in reverse-
  and forward-Z: the second groom's G-buffer and the composited depth bit for
  bit, colour within 1e-3.

Caller-owned device buffers come from the HIP runtime directly
(tests/devmem.py); every result comes from libhairmi.so.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from bevy_hair_amd import plugin, scene
from bevy_hair_amd.groom import synthetic_groom
from tests.branch_cases import _with_ao
from tests.devmem import DeviceArray
from tests.scenes import EMPTY64, make_scene, rel_err

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _points(mesh, n_box=40, n_plane=120):
    """World points: a grid over the groom's bounds grown by 30 %, and a plane
    a little below the groom (float4, w = 1)."""
    lo, hi = mesh.bounds_min.astype(np.float64), mesh.bounds_max.astype(np.float64)
    c, h = 0.5 * (lo + hi), 0.65 * (hi - lo)
    g = [np.linspace(c[k] - h[k], c[k] + h[k], n_box) for k in range(3)]
    box = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3)
    px = np.linspace(c[0] - 1.5 * h[0], c[0] + 1.5 * h[0], n_plane)
    pz = np.linspace(c[2] - 1.5 * h[2], c[2] + 1.5 * h[2], n_plane)
    X, Z = np.meshgrid(px, pz, indexing="ij")
    plane = np.stack([X.ravel(), np.full(X.size, lo[1] - 0.1), Z.ravel()], -1)
    pts = np.concatenate([box, plane])
    return np.concatenate([pts, np.ones((len(pts), 1))], 1).astype(np.float32)


@pytest.mark.parametrize("name,dom_size", [("c1", 512), ("c2", 2048)])
def test_shadow_factors_against_oracle(renderer, name, dom_size):
    sc = make_scene(name)
    w = sc.w
    print(f"\n[gpu compose] shadow receiver, {name}, map {dom_size}^2", flush=True)
    params = plugin.Renderer.make_params(sc.view, sc.groom, sc.mesh, dom_size, sc.dirs)
    r = renderer
    r.targets(w.width, w.height, dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    try:
        r.hair_pass(g, params, r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction))
        light = r.shadow_light(params[1])
        assert light.depth[3] == 1.0
        pts = _points(sc.mesh)
        dom_ptr, _ = r.device_ptr(N.BUF_DOM)
        dev_pts = DeviceArray(pts)
        out = DeviceArray(np.full(len(pts), -1.0, np.float32))
        r.shadow_factors(light, dom_ptr, dom_size, dev_pts.ptr, len(pts), out.ptr)
        r.sync()
        got = out.download()
        dom = r.dom_map()
        dev_pts.free()
        out.free()
    finally:
        g.free()
    L = oracle.lib()
    import ctypes as C

    ref = np.array([L.or_hair_visibility(C.byref(light), dom.ctypes.data, dom_size,
                                         (C.c_float * 3)(*map(float, p[:3]))) for p in pts], np.float32)
    err = np.abs(got - ref)
    shadowed = (ref < 0.99).sum()
    print(f"map {dom_size}: {len(pts)} points, {shadowed} shadowed, max |diff| {err.max():.2e}")
    assert shadowed > 1000 and (ref == 1.0).sum() > 1000
    assert err.max() <= 1e-3
    assert (got >= 0.0).all() and (got <= 1.0 + 1e-6).all()


@pytest.mark.parametrize("projection", ["perspective", "perspective_forward_z"])
def test_two_grooms_composite_through_the_view_depth(renderer, projection):
    W, H, DOM = 480, 270, 512
    mesh = _with_ao(synthetic_groom("straight", 4096, 1))
    view = scene.fit_camera(mesh, W, H, projection=projection)
    reverse = projection == "perspective"
    clear = 0.0 if reverse else 1.0
    radius = 0.5 * float(np.linalg.norm(mesh.bounds_max - mesh.bounds_min))
    right, back = view.world_from_view[:3, 0], view.world_from_view[:3, 2]
    shift = np.eye(4)
    shift[:3, 3] = 0.45 * radius * right + 0.35 * radius * back  # in front of the first groom, half over it
    grooms = [plugin.HairGroom(), plugin.HairGroom(world_from_model=shift)]
    rng = np.random.default_rng(31)
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64)
    points, dirs = scene.point_lights(rng, 2, centre), [scene.key_light()]
    params = [plugin.Renderer.make_params(view, gr, mesh, DOM, dirs) for gr in grooms]
    assert params[0][0].pass_mode[2] == (1 if reverse else 0)
    print(f"\n[gpu compose] two grooms, {projection}", flush=True)
    r = renderer
    tc = DeviceArray(np.zeros((W * H, 4), np.float32))
    td = DeviceArray(np.full(W * H, clear, np.float32))
    r.targets(W, H, DOM)
    r.bind_scene_depth(td.ptr, clear)
    r.set_lights(points, dirs)
    g = r.upload(mesh)
    opts = r.frame_opts(control_points=17, strand_fraction=1.0)
    try:
        for p in params:
            r.hair_pass(g, p, opts)
            gb_last = r.gbuffer()
            r.composite(p[0], tc.ptr, td.ptr)
        r.sync()
        got_c, got_d = tc.download(), td.download()
    finally:
        g.free()
        r.set_scene_depth(None, 0.0)
        tc.free()
        td.free()
    # the oracle: the same two grooms, each raster reading the composited view depth
    pa, n_p, da, n_d = plugin.light_arrays(points, dirs)
    ref_c = np.zeros((W * H, 4), np.float32)
    ref_d = np.full(W * H, clear, np.float32)
    for p in params:
        f = oracle.Frame(mesh, W, H, DOM)
        f.scene_depth, f.scene_clear = ref_d, clear
        f.hair_pass(p, pa, n_p, da, n_d, control_points=17, strand_fraction=1.0, threads=THREADS)
        f.composite(p[0], ref_c, ref_d)
    center, cons, beta = gb_last
    assert np.array_equal(center, f.center), f"{(center != f.center).sum()} centre words of the second groom differ"
    assert np.array_equal(cons, f.cons) and np.array_equal(beta, f.beta)
    # the first groom hides part of the second: without the view depth it covers more
    free = oracle.Frame(mesh, W, H, DOM)
    free.scene_clear = clear
    free.hair_pass(params[1], pa, n_p, da, n_d, control_points=17, strand_fraction=1.0, threads=THREADS)
    hidden = int((free.center != EMPTY64).sum() - (f.center != EMPTY64).sum())
    print(f"{projection}: second groom loses {hidden} centre pixels to the first; "
          f"{int((got_c[:, 3] > 0).sum())} composited pixels")
    assert hidden > 200
    assert np.array_equal(got_d, ref_d), f"{(got_d != ref_d).sum()} composited depths differ"
    assert np.array_equal(got_c[:, 3], ref_c[:, 3]), "composited alpha differs"
    e = rel_err(got_c[:, :3], ref_c[:, :3])
    assert e.max() <= 1e-3, e.max()
