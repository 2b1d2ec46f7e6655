"""hm_autotune on a small scene: the choice it returns lies in its search
space (camera tile 0 / 16 / 32 x 64 .. 2048 records per workgroup, DOM tile
0 / 16 / 32 x 1 / 2 / 4 binning workgroups per CU x persistent tile
workgroups 0 / 1 / 2, frame on one or two streams), the context keeps it, and
a frame with it gives the same bytes as a frame with the direct kernels on
one stream (the results never depend on the launch choices)."""
from __future__ import annotations

import numpy as np
import pytest

from tests.scenes import make_scene

pytestmark = pytest.mark.gpu


def _outputs(r):
    center, cons, beta = r.gbuffer()
    shaded, _ = r.shaded()
    filtered, _ = r.filtered()
    return [center, cons, beta, r.dom_map(), shaded, filtered]


def test_autotune_choice_and_results(renderer):
    sc = make_scene("c1")
    w = sc.w
    print("\n[gpu autotune] c1", flush=True)
    r = renderer
    keep = r.get_tuning()
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    try:
        r.set_tuning(cam_tile=0, dom_tile=0, frame_serial=1)
        r.hair_pass(g, sc.params, opts)
        want = [a.copy() for a in _outputs(r)]
        t = r.autotune(g, sc.params, opts, reps=1)
        assert t == r.get_tuning()
        assert t["cam_tile"] in (0, 16, 32) and t["dom_tile"] in (0, 16, 32)
        if t["cam_tile"]:
            assert t["cam_chunk"] in (64, 128, 256, 512, 1024, 2048)
        assert t["bin_blocks_per_cu"] in (1, 2, 4)
        assert t["dom_busy_blocks_per_cu"] in (0, 1, 2) and t["frame_serial"] in (0, 1) and t["filter_tiles"] in (0, 1)
        assert all(f in (0, 1, 2, 3) for f in t["raster_flags"])
        assert all(b in (4, 8, 16, 32) for b in t["raster_blocks_per_cu"])
        r.hair_pass(g, sc.params, opts)
        for i, (a, b) in enumerate(zip(_outputs(r), want)):
            assert np.array_equal(a, b), f"output {i} differs with the tuned launch choices {t}"
    finally:
        r.set_tuning(**keep)
        g.free()
