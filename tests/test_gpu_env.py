"""The environment probe, rect lights and per-wave light culling on the HIP
path against the oracle (shade.wesl:396-427, 503-518, 448-501, 586).

- a groom with an AO ramp under a key light, 400 short-range point and spot
  lights scattered around it (most waves reach only a few of them, so the
  culling drops most of the list), three rect lights and a rotated studio
  probe: the G-buffer and DOM the GPU wrote are shaded again by the oracle,
  colours within 1e-3;
- culling changes nothing: 600 more lights whose range stops short of the
  groom leave every colour bit for bit as it was;
- the probe term is additive: removing the probe changes each colour by the
  oracle's probe x AO term alone;
- the kernels with and without per-wave culling give the same bits.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from bevy_hair_amd import plugin, scene
from bevy_hair_amd.groom import synthetic_groom
from tests.branch_cases import _with_ao
from tests.scenes import rel_err

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
W, H, DOM = 640, 360, 512
RGB_TOL = 1e-3


def far_lights(rng, n, centre, radius):
    """Lights 3-5 radii from the groom's centre whose range ends a radius
    short of it: none of them reaches a pixel."""
    out = []
    for _ in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(3.0, 5.0) * radius
        rng_ = dist - 2.0 * radius
        out.append({"position_radius": [*(centre + d * dist).tolist(), 0.0],
                    "color_inverse_square_range": [1.0, 1.0, 1.0, 1.0 / (rng_ * rng_)],
                    "light_custom_data": [0.0, 0.0, 0.0, 0.0], "flags": 0, "kind": 0})
    return out


@pytest.fixture(scope="module")
def setup(renderer):
    mesh = _with_ao(synthetic_groom("straight", 4096, 1))
    view = scene.fit_camera(mesh, W, H)
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64)
    radius = 0.5 * float(np.linalg.norm(mesh.bounds_max - mesh.bounds_min))
    rng = np.random.default_rng(77)
    points = scene.scattered_lights(rng, 400, centre, radius, 0.15, 0.6)
    dirs = [scene.key_light()]
    rects = scene.studio_rects(centre, radius)
    q = np.array([0.1, 0.7, -0.2, 0.68])
    env = scene.studio_environment(64, 7, 16, intensity=0.8, rotation=tuple((q / np.linalg.norm(q)).tolist()))
    groom = plugin.HairGroom()
    params = plugin.Renderer.make_params(view, groom, mesh, DOM, dirs)
    assert params[0].flags[2] == 1  # AO on
    r = renderer
    r.targets(W, H, DOM)
    r.set_scene_depth(None, 0.0)
    g = r.upload(mesh)
    opts = r.frame_opts(control_points=17, strand_fraction=1.0)

    def frame(pts, with_env=True):
        r.set_lights(pts, dirs)
        r.set_rect_lights(rects)
        r.set_environment(env if with_env else None)
        r.hair_pass(g, params, opts)
        r.sync()
        return r.shaded()

    try:
        print(f"\n[gpu env] {len(points)} short lights, {len(rects)} rects, probe 64^2 x 7", flush=True)
        out = {"base": frame(points)}
        out["gb"], out["dom"] = r.gbuffer(), r.dom_map()
        print("[gpu env] with far lights", flush=True)
        out["far"] = frame(points + far_lights(rng, 600, centre, radius))
        print("[gpu env] without the probe", flush=True)
        out["no_env"] = frame(points, with_env=False)
        # 40 lights: the frame culls them per wave (some ranges end inside
        # the groom's box); hm_shade on its own, with no groom box and fewer
        # than 64 lights, does not
        print("[gpu env] 40 lights, culled and not", flush=True)
        out["few_culled"] = frame(points[:40])
        r.shade(params[0], 1.0)
        r.sync()
        out["few_plain"] = r.shaded()
    finally:
        r.set_rect_lights([])
        r.set_environment(None)
        g.free()
    out.update(mesh=mesh, params=params, points=points, dirs=dirs, rects=rects, env=env)
    return out


def oracle_shade(s, with_env=True):
    f = oracle.Frame(s["mesh"], W, H, DOM)
    f.center[:], f.cons[:], f.beta[:] = s["gb"]
    f.dom[:] = s["dom"]
    ra = plugin.rect_arrays(s["rects"])
    f.rects = ra
    desc = plugin.EnvProbeDesc(s["env"]) if with_env else None
    f.env = desc.desc if desc else None
    pa, n_p, da, n_d = plugin.light_arrays(s["points"], s["dirs"])
    f.shade(s["params"][0], pa, n_p, da, n_d, threads=THREADS)
    return f


def test_probe_rects_and_short_lights_against_oracle(setup):
    f = oracle_shade(setup)
    shaded, depth = setup["base"]
    hit = f.shaded[:, 3] > 0
    assert hit.sum() > 5000
    assert np.array_equal(depth, f.shaded_depth)
    e = rel_err(shaded, f.shaded)
    print(f"shade max rel err {e.max():.3e} over {int(hit.sum())} pixels")
    assert e.max() <= RGB_TOL


def test_lights_out_of_range_change_no_bit(setup):
    assert np.array_equal(setup["far"][0], setup["base"][0])
    assert np.array_equal(setup["far"][1], setup["base"][1])


def test_probe_term_is_additive(setup):
    with_env = oracle_shade(setup)
    without = oracle_shade(setup, with_env=False)
    probe_term = with_env.shaded[:, :3].astype(np.float64) - without.shaded[:, :3]
    hit = with_env.shaded[:, 3] > 0
    assert np.abs(probe_term[hit]).min(axis=1).max() > 0.0
    gpu_term = setup["base"][0][:, :3].astype(np.float64) - setup["no_env"][0][:, :3]
    e = np.abs(gpu_term - probe_term) / np.maximum(np.abs(with_env.shaded[:, :3]), 1e-3)
    print(f"probe term max rel err {e.max():.3e}")
    assert e.max() <= 2 * RGB_TOL


def test_culled_and_unculled_shading_agree_bit_for_bit(setup):
    assert np.array_equal(setup["few_culled"][0], setup["few_plain"][0])
    assert np.array_equal(setup["few_culled"][1], setup["few_plain"][1])
