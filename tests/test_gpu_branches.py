"""The branches of hair_pass outside the measured workloads, HIP path against
the oracle through hm_frame: spot lights, several directional lights, no
shadow caster, forward-Z and orthographic cameras with scene depth, and the
HairGroom toggles lod / ambient_occlusion / deep_opacity off (the cases of
tests/branch_cases.py). Binned and direct raster kernels both run each case.
G-buffer words and DOM bit for bit; shaded and filtered RGB within 1e-3
relative; alpha and depths bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from tests.branch_cases import CASE_NAMES, DOM, HEIGHT, WIDTH, make_case, oracle_frame
from tests.scenes import EMPTY64, rel_err

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-3
TILINGS = {"binned": dict(cam_tile=16, dom_tile=16), "direct": dict(cam_tile=0, dom_tile=0)}


@pytest.fixture(scope="module", params=CASE_NAMES)
def case(request):
    c = make_case(request.param)
    return c, oracle_frame(c)


@pytest.mark.parametrize("tiling", list(TILINGS))
def test_branch_case_against_oracle(renderer, case, tiling):
    c, ref = case
    print(f"\n[gpu branches] case={c.name} tiling={tiling} reverse={c.params[0].pass_mode[2]}", flush=True)
    r = renderer
    keep = r.get_tuning()
    g = None
    try:
        r.targets(WIDTH, HEIGHT, DOM)
        r.set_scene_depth(c.scene_depth, c.scene_clear)
        r.set_lights(c.points, c.dirs)
        r.set_tuning(**TILINGS[tiling])
        g = r.upload(c.mesh)
        r.hair_pass(g, c.params, r.frame_opts(control_points=c.control_points, strand_fraction=c.strand_fraction))
        center, cons, beta = r.gbuffer()
        assert (center != EMPTY64).sum() > 1000, "nothing rasterized"
        assert np.array_equal(center, ref.center), f"{(center != ref.center).sum()} centre words differ"
        assert np.array_equal(cons, ref.cons), f"{(cons != ref.cons).sum()} conservative words differ"
        assert np.array_equal(beta, ref.beta), f"{(beta != ref.beta).sum()} beta words differ"
        if c.params[0].flags[3]:
            dmap = r.dom_map()
            assert np.array_equal(dmap, ref.dom), f"{(dmap != ref.dom).sum()} DOM words differ"
        shaded, depth = r.shaded()
        assert np.array_equal(depth, ref.shaded_depth)
        e = rel_err(shaded, ref.shaded)
        filtered, fdepth = r.filtered()
        assert np.array_equal(fdepth, ref.filtered_depth)
        assert np.array_equal(filtered[:, 3], ref.filtered[:, 3]), "filter alpha differs"
        ef = rel_err(filtered[:, :3], ref.filtered[:, :3])
        print(f"{c.name}/{tiling}: shade max rel err {e.max():.2e}, filter {ef.max():.2e}")
        assert e.max() <= RGB_TOL and ef.max() <= RGB_TOL
    finally:
        if g is not None:
            g.free()
        r.set_scene_depth(None, 0.0)
        r.set_tuning(**keep)
