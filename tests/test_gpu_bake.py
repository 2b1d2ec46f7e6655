"""A groom from the baker (hm_bake) rendered by the HIP path against the
oracle: 3000 wavy strands of 48 points at the reference assets' centimetre
scale, baked (clustering, coherence splits, style, ambient occlusion), then
one hm_frame at 480 x 270 with a 512^2 DOM. G-buffer words (their AO field
now carries the baked occlusion) and DOM bit for bit; shade and filter within
1e-3."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from bevy_hair_amd import plugin, scene
from bevy_hair_amd.bake import bake_strands
from bevy_hair_amd.groom import synth_strands
from tests.scenes import EMPTY64, rel_err

pytestmark = pytest.mark.gpu


def test_baked_groom_frame_against_oracle(renderer):
    raw = synth_strands("wavy", 3000, 5) * 30.0  # unit-sphere scalp -> 30 units, as the example grooms
    points = raw.reshape(-1, 3).astype(np.float32)
    strands = np.stack([np.arange(len(raw)) * raw.shape[1], np.full(len(raw), raw.shape[1])], 1).astype(np.uint32)
    mesh = bake_strands(points, strands)
    assert int(mesh.bundles[:, 2].sum()) == 3000
    ao = mesh.layers[:, 3]
    assert ao.min() < 0.9 < ao.max(), "the baked occlusion should vary"
    W, H, DOM = 480, 270, 512
    view = scene.fit_camera(mesh, W, H)
    rng = np.random.default_rng(12)
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64)
    points_l, dirs = scene.point_lights(rng, 2, centre, box=60.0, light_range=80.0), [scene.key_light()]
    params = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, DOM, dirs)
    print(f"\n[gpu bake] {mesh.bundle_count} baked bundles", flush=True)
    r = renderer
    r.targets(W, H, DOM)
    r.set_scene_depth(None, 0.0)
    r.set_lights(points_l, dirs)
    g = r.upload(mesh)
    try:
        r.hair_pass(g, params, r.frame_opts())
        center, cons, beta = r.gbuffer()
        dmap = r.dom_map()
        shaded, sdepth = r.shaded()
        filtered, fdepth = r.filtered()
    finally:
        g.free()
    ref = oracle.Frame(mesh, W, H, DOM)
    pa, n_p, da, n_d = plugin.light_arrays(points_l, dirs)
    ref.hair_pass(params, pa, n_p, da, n_d, threads=8)
    hit = center != EMPTY64
    assert hit.sum() > 2000
    assert len(np.unique(center[hit] & np.uint64(63))) > 4, "AO codes should vary"
    assert np.array_equal(center, ref.center) and np.array_equal(cons, ref.cons) and np.array_equal(beta, ref.beta)
    assert np.array_equal(dmap, ref.dom)
    assert np.array_equal(sdepth, ref.shaded_depth) and rel_err(shaded, ref.shaded).max() <= 1e-3
    assert np.array_equal(filtered[:, 3], ref.filtered[:, 3]) and rel_err(filtered[:, :3], ref.filtered[:, :3]).max() <= 1e-3
