"""The batched multi-view frame (hm_batch_*): a captured HIP graph of several
views of several grooms at mixed sizes must give, on every replay, exactly the
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


@pytest.mark.parametrize("mlp", [False, True])
def test_batch_replays_equal_one_view_at_a_time(renderer, mlp):
    print(f"\n[gpu batch] 6 views, mlp={mlp}", flush=True)
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
