"""The frame's filter by 16 x 16 screen tiles with its taps in LDS
(hm_tuning.filter_tiles = 1) against the filter over the camera tiles' pixel
lists with its taps from memory (filter_tiles = 0, the default): the filtered
colour, alpha and depth are the same bit for bit, on views whose size is not
a multiple of the tile, with tiles split into several record chunks, and with
long segments whose pixels reach the lists after the tiles."""
from __future__ import annotations

import numpy as np
import pytest

from tests.scenes import make_scene

pytestmark = pytest.mark.gpu

CASES = [
    # scene, camera tile, records per camera tile workgroup, tiles a segment's box may span
    ("c1", 16, 2048, 8),
    ("c1", 32, 128, 8),        # most tiles split into chunks
    ("c1", 16, 2048, 1),       # most segments long
    ("small_lod", 32, 2048, 8),
    ("dense_dom", 32, 512, 8),
    ("c2", 32, 2048, 8),       # the measured 1080p frame
]


@pytest.mark.parametrize("name,cam_tile,chunk,box_limit", CASES)
def test_tile_filter_equals_list_filter(renderer, name, cam_tile, chunk, box_limit):
    sc = make_scene(name)
    w = sc.w
    print(f"\n[gpu filter tiles] {name} {w.width}x{w.height} cam_tile={cam_tile} chunk={chunk} "
          f"box_limit={box_limit}", flush=True)
    r = renderer
    keep = r.get_tuning()
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    try:
        r.set_tuning(cam_tile=cam_tile, cam_chunk=chunk, bin_box_limit=box_limit, filter_tiles=0)
        r.hair_pass(g, sc.params, opts)
        st = r.bin_status()
        want = [a.copy() for a in r.filtered()]
        assert st["binned"] and (want[0][:, 3] > 0).sum() > 100, "the view shows no filtered hair"
        if box_limit == 1:
            assert st["long"] and st["long_count"] > 0
        r.set_tuning(filter_tiles=1)
        for k in range(2):
            r.hair_pass(g, sc.params, opts)
            got = r.filtered()
            assert np.array_equal(got[0], want[0]), f"filtered colour differs (frame {k})"
            assert np.array_equal(got[1], want[1]), f"filtered depth differs (frame {k})"
    finally:
        r.set_tuning(**keep)
        g.free()
