"""The batched multi-view frame (hm_batch_*): a captured HIP graph of several
views of several grooms at mixed sizes must give, on every replay and for
every lane count (views dealt to one to three concurrent chains), exactly the
images the same views give rendered one at a time with hm_frame."""
from __future__ import annotations

import numpy as np
import pytest

from bevy_hair_amd import mlp_train, plugin, scene
from bevy_hair_amd.groom import synthetic_groom

pytestmark = pytest.mark.gpu

GROOMS = [("straight", 4096, 1, 17), ("wavy", 2048, 11, 33)]
SIZES = [(640, 360), (320, 180), (800, 450)]


def _views(renderer, meshes, grooms, dirs, mlp):
    gpu = [renderer.upload(m) for m in meshes]
    views = []
    for v in range(6):
        g = v % len(meshes)
        w, h = SIZES[v % len(SIZES)]
        cam = scene.fit_camera(meshes[g], w, h, heading=(150.0 - 40.0 * v, 0.0, 40.0 + 30.0 * v))
        dom = 512 if v % 2 else 256
        params = plugin.Renderer.make_params(cam, grooms[g], meshes[g], dom, dirs)
        opts = renderer.frame_opts(control_points=GROOMS[g][3], strand_fraction=1.0, mlp=mlp and v % 3 != 1,
                                   mlp_tolerance=0.0)
        views.append((gpu[g], params, w, h, dom, opts))
    return gpu, views


@pytest.mark.parametrize("lanes", [1, 2, 3])
@pytest.mark.parametrize("mlp", [False, True])
def test_batch_replays_equal_one_view_at_a_time(renderer, mlp, lanes):
    print(f"\n[gpu batch] 6 views, mlp={mlp}, lanes={lanes}", flush=True)
    keep = renderer.get_tuning()["batch_lanes"]
    renderer.set_tuning(batch_lanes=lanes)
    mlp_train.install(renderer)
    meshes = [synthetic_groom(kind, n, seed) for kind, n, seed, _ in GROOMS]
    grooms = [plugin.HairGroom(filter=True), plugin.HairGroom(filter=False)]
    rng = np.random.default_rng(3)
    points = scene.point_lights(rng, 4, 0.5 * (meshes[0].bounds_min + meshes[0].bounds_max))
    dirs = [scene.key_light()]
    renderer.set_scene_depth(None, 0.0)
    renderer.set_lights(points, dirs)
    gpu, views = _views(renderer, meshes, grooms, dirs, mlp)
    batch = renderer.batch(views)
    try:
        batch.run(2)
        first = [batch.read(i) for i in range(len(views))]
        batch.run(1)
        for i in range(len(views)):
            rgba, depth = batch.read(i)
            assert np.array_equal(rgba, first[i][0]) and np.array_equal(depth, first[i][1]), f"replay differs, view {i}"
        for i, (g, params, w, h, dom, opts) in enumerate(views):
            renderer.targets(w, h, dom)
            renderer.hair_pass(g, params, opts)
            img, depth = renderer.filtered() if params[0].flags[1] else renderer.shaded()
            assert (img[:, 3] > 0).sum() > 100, f"view {i} shows no hair"
            assert np.array_equal(first[i][0], img), f"view {i} colour differs from hm_frame"
            assert np.array_equal(first[i][1], depth), f"view {i} depth differs from hm_frame"
        ms = batch.time(3) / 3
        print(f"6 views: {ms:.3f} ms per replay")
    finally:
        batch.free()
        for g in gpu:
            g.free()
        renderer.set_tuning(batch_lanes=keep)


@pytest.mark.parametrize("lanes", [2, 3])
def test_batch_lanes_with_per_view_lod(renderer, lanes):
    """Views of the same groom with the reference's screen-space LOD (no
    override) at distances and sizes that give each a different LOD record,
    dealt so that several lanes render one groom at the same time: each lane
    keeps its own LOD records and strand refs, and every replay equals the
    views rendered one at a time."""
    print(f"\n[gpu batch] per-view LOD, lanes={lanes}", flush=True)
    keep = renderer.get_tuning()["batch_lanes"]
    renderer.set_tuning(batch_lanes=lanes)
    meshes = [synthetic_groom("wavy", 2048, 11), synthetic_groom("straight", 3072, 4)]
    grooms = [plugin.HairGroom(filter=True), plugin.HairGroom(filter=True)]
    dirs = [scene.key_light()]
    renderer.set_scene_depth(None, 0.0)
    renderer.set_lights(scene.point_lights(np.random.default_rng(9), 3,
                                           0.5 * (meshes[0].bounds_min + meshes[0].bounds_max)), dirs)
    gpu = [renderer.upload(m) for m in meshes]
    views = []
    for v in range(8):
        gi = (v // 3) % 2  # runs of three views per groom: every lane count puts one groom on several lanes
        w, h = (640, 360) if v % 2 else (320, 180)
        cam = scene.fit_camera(meshes[gi], w, h, heading=(120.0 - 35.0 * v, 0.0, 50.0), margin=1.1 + 1.7 * v)
        params = plugin.Renderer.make_params(cam, grooms[gi], meshes[gi], 256, dirs)
        views.append((gpu[gi], params, w, h, 256, renderer.frame_opts()))
    batch = None
    try:
        lods = set()
        want = []
        for g, params, w, h, dom, opts in views:
            renderer.targets(w, h, dom)
            renderer.hair_pass(g, params, opts)
            want.append([a.copy() for a in renderer.filtered()])
            lods.add(renderer.lod_records(g, 0).tobytes())
        assert len(lods) >= 4, "the views should not share their LOD records"
        batch = renderer.batch(views)
        for rep in range(2):
            batch.run(1)
            for i in range(len(views)):
                rgba, depth = batch.read(i)
                assert np.array_equal(rgba, want[i][0]), f"replay {rep}: view {i} colour differs from hm_frame"
                assert np.array_equal(depth, want[i][1]), f"replay {rep}: view {i} depth differs from hm_frame"
    finally:
        if batch is not None:
            batch.free()
        for g in gpu:
            g.free()
        renderer.set_tuning(batch_lanes=keep)


def test_batch_with_scene_depth(renderer):
    """With a scene depth set the views stay in one chain (the depth belongs to
    the context's target): four views of one size in front of a checkerboard
    depth equal hm_frame with the same depth, on two replays."""
    print("\n[gpu batch] 4 views over a scene depth, batch_lanes=2", flush=True)
    keep = renderer.get_tuning()["batch_lanes"]
    renderer.set_tuning(batch_lanes=2)
    mesh = synthetic_groom("straight", 4096, 1)
    groom = plugin.HairGroom(filter=True)
    rng = np.random.default_rng(5)
    renderer.set_lights(scene.point_lights(rng, 3, 0.5 * (mesh.bounds_min + mesh.bounds_max)), [scene.key_light()])
    w, h, dom = 480, 270, 256
    g = renderer.upload(mesh)
    views, cams = [], []
    for v in range(4):
        cam = scene.fit_camera(mesh, w, h, heading=(120.0 - 50.0 * v, 0.0, 60.0 + 20.0 * v))
        params = plugin.Renderer.make_params(cam, groom, mesh, dom, [scene.key_light()])
        views.append((g, params, w, h, dom, renderer.frame_opts(control_points=17, strand_fraction=1.0)))
        cams.append(cam)
    # every other 24-pixel square at the clip depth of the groom's centre in the first view
    c = np.append(0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64), 1.0)
    clip = cams[0].clip_from_world @ c
    yy, xx = np.mgrid[0:h, 0:w]
    depth_map = np.where(((xx // 24) + (yy // 24)) % 2 == 0, np.float32(clip[2] / clip[3]),
                         np.float32(0.0)).astype(np.float32).ravel()
    renderer.targets(w, h, dom)
    renderer.set_scene_depth(depth_map, 0.0)
    batch = None
    try:
        batch = renderer.batch(views)
        batch.run(1)
        first = [batch.read(i) for i in range(len(views))]
        batch.run(1)
        occluded = 0
        for i, (gg, params, _, _, _, opts) in enumerate(views):
            rgba, d = batch.read(i)
            assert np.array_equal(rgba, first[i][0]) and np.array_equal(d, first[i][1]), f"replay differs, view {i}"
            renderer.hair_pass(gg, params, opts)
            img, dd = renderer.filtered()
            assert np.array_equal(rgba, img) and np.array_equal(d, dd), f"view {i} differs from hm_frame"
            renderer.set_scene_depth(None, 0.0)
            renderer.hair_pass(gg, params, opts)
            occluded += int((renderer.filtered()[0][:, 3] > 0).sum() - (img[:, 3] > 0).sum())
            renderer.set_scene_depth(depth_map, 0.0)
        assert occluded > 0, "the scene depth hid no hair pixel"
    finally:
        if batch is not None:
            batch.free()
        renderer.set_scene_depth(None, 0.0)
        g.free()
        renderer.set_tuning(batch_lanes=keep)


def test_c5_batch_against_frames_and_oracle(renderer):
    """The measured batch: 24 views over the 20k x 16, 100k x 64 and 150k x 126
    grooms at 1280x720 / 1920x1080 / 2560x1440, DOM 2048^2, learned shading
    with the trained-range fallback. Replays are identical; every view equals
    hm_frame; for one view of each groom the G-buffer and DOM are bit-exact
    against the oracle, the same view shaded by the plain method is within
    1e-3 of the oracle, and the batch's learned shading is within 2 % (mean
    relative RGB) of it."""
    import os

    import oracle
    from bevy_hair_amd import workload as W
    from tests.scenes import EMPTY64, rel_err

    threads = min(16, os.cpu_count() or 1)
    print("\n[gpu batch] c5: 24 views, three grooms, mixed sizes", flush=True)
    bs = W.make_batch_scene()
    mlp_train.install(renderer)
    renderer.set_scene_depth(None, 0.0)
    renderer.set_lights(bs.points, bs.dirs)
    gpu = [renderer.upload(m) for m in bs.meshes]
    views = [(gpu[gi], params, wd, ht, W.BATCH_DOM,
              renderer.frame_opts(control_points=ncp, strand_fraction=1.0, mlp=True, mlp_tolerance=0.0))
             for gi, wd, ht, _, params, ncp in bs.views]
    batch = renderer.batch(views)
    pa, n_p, da, n_d = plugin.light_arrays(bs.points, bs.dirs)
    try:
        batch.run(2)
        first = [batch.read(i) for i in range(len(views))]
        batch.run(1)
        checked = set()
        for i, (g, params, w, h, dom, opts) in enumerate(views):
            rgba, depth = batch.read(i)
            assert np.array_equal(rgba, first[i][0]) and np.array_equal(depth, first[i][1]), f"replay differs, view {i}"
            renderer.targets(w, h, dom)
            renderer.hair_pass(g, params, opts)
            img, d = renderer.filtered()
            assert (img[:, 3] > 0).sum() > 1000, f"view {i} shows no hair"
            assert np.array_equal(first[i][0], img), f"view {i} colour differs from hm_frame"
            assert np.array_equal(first[i][1], d), f"view {i} depth differs from hm_frame"
            gi = bs.views[i][0]
            if gi in checked:
                continue
            checked.add(gi)
            print(f"[gpu batch] c5 view {i}: groom {gi} {w}x{h} against the oracle", flush=True)
            mlp_rgb = renderer.shaded()[0].copy()
            center, cons, beta = renderer.gbuffer()
            dmap = renderer.dom_map()
            ref = oracle.Frame(bs.meshes[gi], w, h, dom)
            cap = max(1, bs.meshes[gi].strand_count)
            lod = ref.lod(params[0], opts.camera_lod.control_points, opts.camera_lod.strand_fraction)
            refs, _ = ref.scan(params[0], lod, cap)
            ref.clear()
            ref.raster(params[0], refs, threads=threads)
            assert (center != EMPTY64).sum() > 1000
            assert np.array_equal(center, ref.center), f"view {i}: {(center != ref.center).sum()} centre words differ"
            assert np.array_equal(cons, ref.cons) and np.array_equal(beta, ref.beta), f"view {i}: G-buffer differs"
            dlod = ref.lod(params[1])
            drefs, _ = ref.scan(params[1], dlod, cap)
            ref.clear_dom()
            ref.raster(params[1], drefs, threads=threads)
            ref.raster(params[2], drefs, threads=threads)
            assert np.array_equal(dmap, ref.dom), f"view {i}: {(dmap != ref.dom).sum()} DOM words differ"
            plain_opts = renderer.frame_opts(control_points=opts.camera_lod.control_points, strand_fraction=1.0)
            renderer.hair_pass(g, params, plain_opts)
            plain = renderer.shaded()[0]
            ref.shade(params[0], pa, n_p, da, n_d, threads=threads)
            e = rel_err(plain, ref.shaded)
            assert e.max() <= 1e-3, f"view {i}: plain shading max rel err {e.max():.3e}"
            hit = ref.shaded_depth > 0
            num = np.abs(mlp_rgb[hit, :3].astype(np.float64) - plain[hit, :3]).sum(axis=1)
            den = np.maximum(np.abs(plain[hit, :3].astype(np.float64)).sum(axis=1), 1e-3)
            mean_err = float((num / den).mean())
            print(f"[gpu batch] c5 view {i}: plain max rel err {e.max():.2e}, learned mean rel err {mean_err:.4f}")
            assert mean_err <= 0.02
        assert checked == {0, 1, 2}
    finally:
        batch.free()
        for g in gpu:
            g.free()
