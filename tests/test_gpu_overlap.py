"""The frame with the DOM pass on a second stream beside the camera pass
(hm_tuning.frame_serial = 0, the default) against the same frame with the
passes one after the other on one stream (frame_serial = 1): every output is
the same bit for bit, frame after frame, including when the camera pass
falls back to its direct kernels (nothing for the DOM pass to reuse) and
when the DOM pass's work lists differ from the camera's (small_lod)."""
from __future__ import annotations

import numpy as np
import pytest

from tests.scenes import make_scene

pytestmark = pytest.mark.gpu


def _outputs(r):
    center, cons, beta = r.gbuffer()
    shaded, shaded_depth = r.shaded()
    filtered, filtered_depth = r.filtered()
    return [center, cons, beta, r.dom_map(), shaded, shaded_depth, filtered, filtered_depth]


@pytest.mark.parametrize("name,cam_tile", [("c1", 16), ("c1", 0), ("small_lod", 16), ("dense_dom", 32)])
def test_two_stream_frame_equals_one_stream(renderer, name, cam_tile):
    sc = make_scene(name)
    w = sc.w
    print(f"\n[gpu overlap] {name} cam_tile={cam_tile}", flush=True)
    r = renderer
    keep = r.get_tuning()
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    try:
        r.set_tuning(cam_tile=cam_tile, frame_serial=1)
        r.hair_pass(g, sc.params, opts)
        want = [a.copy() for a in _outputs(r)]
        assert (want[0] != np.uint64(2**64 - 1)).any(), "the view shows no hair"
        r.set_tuning(frame_serial=0)
        for k in range(3):
            r.hair_pass(g, sc.params, opts)
            got = _outputs(r)
            for i, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), f"output {i} differs on two streams (frame {k})"
    finally:
        r.set_tuning(**keep)
        g.free()
