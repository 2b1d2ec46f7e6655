"""The branch cases of tests/branch_cases.py exercise what they claim to, on
the oracle (CPU): the spot lights reach part of the groom with pixels in the
cone falloff, the DOM follows the first shadow caster, the toggles change the
words they control. The GPU parity of the same cases is test_gpu_branches.py."""
from __future__ import annotations

import numpy as np
import pytest

from tests.branch_cases import DOM, make_case, oracle_frame
from tests.scenes import EMPTY64


@pytest.fixture(scope="module")
def frames():
    return {n: (make_case(n), oracle_frame(make_case(n), threads=4))
            for n in ("spot", "directionals", "ao_off", "dom_off", "lod_off", "forward_z")}


def test_spot_case_lights_part_of_the_groom(frames):
    c, ref = frames["spot"]
    no_spot = make_case("spot")
    no_spot.points = no_spot.points[:1]
    base = oracle_frame(no_spot, threads=4)
    hit = ref.center != EMPTY64
    share = (ref.shaded[hit, :3] - base.shaded[hit, :3]).sum(1) / np.maximum(ref.shaded[hit, :3].sum(1), 1e-6)
    assert (share > 0.5).mean() > 0.25, "spot lights light too little of the groom"
    assert (share < 0.05).mean() > 0.25, "spot lights light everything: no cone edge"
    assert ((share > 0.05) & (share < 0.5)).sum() > 100, "no pixels in the cone falloff"
    # both signs of the stored direction's y
    assert {p["flags"] for p in c.points[1:]} == {0, 2}


def test_dom_follows_first_shadow_caster(frames):
    c, _ = frames["directionals"]
    key = np.array(c.dirs[1]["direction_to_light"][:3])
    fwd = np.array(c.params[1].light_forward[:3])
    assert not c.dirs[0]["flags"] and c.dirs[1]["flags"] and c.dirs[2]["flags"]
    assert np.allclose(fwd, -key / np.linalg.norm(key), atol=1e-6)


def test_ao_off_writes_full_ao(frames):
    _, ref = frames["ao_off"]
    hit = ref.center != EMPTY64
    assert ((ref.center[hit] & np.uint64(63)) == 63).all()
    _, on = frames["directionals"]
    assert ((on.center[on.center != EMPTY64] & np.uint64(63)) != 63).any()


def test_dom_off_skips_the_map(frames):
    c, ref = frames["dom_off"]
    assert c.params[0].flags[3] == 0
    assert (ref.dom == 0).all()  # the oracle's map stays as allocated: no DOM pass ran


def test_lod_off_full_strands_at_127_points(frames):
    c, ref = frames["lod_off"]
    lod = ref.lod(c.params[0])
    alive = lod[:, 0] > 0
    assert alive.any()
    assert (lod[alive, 0] == lod[alive, 2]).all() and (lod[alive, 1] == 127).all()


def test_forward_z_scene_depth_occludes(frames):
    c, ref = frames["forward_z"]
    assert c.params[0].pass_mode[2] == 0  # depth_is_reverse is false
    free = make_case("forward_z")
    free.scene_depth = None
    unoccluded = oracle_frame(free, threads=4)
    assert (unoccluded.center != EMPTY64).sum() > (ref.center != EMPTY64).sum() + 100
    assert DOM == 512
