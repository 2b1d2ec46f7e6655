"""Full-size workloads against the oracle (C2: 1080p, 100k x 64; C3: 2160p,
60k x 126 curly, 1 + 128 lights, LOD strand fractions 1, 1/2, 1/4).

The oracle runs multithreaded (atomic compare-and-swap for the raster, rows
for shade and filter), so every output is compared at full size:
- LOD records and the strand work list, exactly;
- the binned G-buffer and DOM bit for bit against the oracle, and against the
  direct kernels (the two GPU paths share no raster code past the walk);
- the DOM invariant: the counts sum to the number of depth-pass writes;
- the whole shaded and filtered frame within 1e-3 relative RGB (alpha and
  depths bit for bit), each oracle stage fed the GPU's previous stage.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from tests.scenes import EMPTY64, make_scene, rel_err

pytestmark = pytest.mark.gpu

HEAVY = [("c2", None), ("c3", 1.0), ("c3", 0.5), ("c3", 0.25)]
RGB_TOL = 1e-3
BIN_MAX_TILES = 16384  # hm_kernels.h: the binned passes take at most this many tiles
THREADS = min(16, os.cpu_count() or 1)


def binned_tile(width: int, height: int) -> int:
    """The smaller camera tile whose grid the binned pass takes."""
    for t in (16, 32):
        if ((width + t - 1) // t) * ((height + t - 1) // t) <= BIN_MAX_TILES:
            return t
    raise AssertionError("no binned tile size fits")


@pytest.fixture(scope="module", params=HEAVY, ids=lambda p: f"{p[0]}-{p[1]}")
def heavy(request, renderer):
    name, frac = request.param
    sc = make_scene(name)
    w = sc.w
    frac = w.strand_fraction if frac is None else frac
    tile = binned_tile(w.width, w.height)
    print(f"\n[gpu heavy] {name} strand fraction {frac} {w.width}x{w.height} dom={w.dom_size} cam_tile={tile}",
          flush=True)
    r = renderer
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    cam, dom1, dom2 = sc.params
    keep = r.get_tuning()
    out = {"scene": sc, "frac": frac}
    try:
        r.set_tuning(cam_tile=tile, dom_tile=16)
        r.lod(g, cam, N.PASS_CAMERA, w.control_points, frac)
        r.scan(g, cam, N.PASS_CAMERA)
        out["cam_lod"] = r.lod_records(g, N.PASS_CAMERA)
        out["cam_refs"] = r.refs(g, N.PASS_CAMERA)
        r.clear()
        r.raster(g, cam, N.PASS_CAMERA)
        out["gb_direct"] = tuple(a.copy() for a in r.gbuffer())
        r.gbuffer_pass(g, cam)
        out["gb"] = r.gbuffer()
        out["cam_status"] = r.bin_status(0)
        r.lod(g, dom1, N.PASS_DOM)
        r.scan(g, dom1, N.PASS_DOM)
        out["dom_lod"] = r.lod_records(g, N.PASS_DOM)
        out["dom_refs"] = r.refs(g, N.PASS_DOM)
        out["dom_counts1"] = r.raster_count(g, dom1, N.PASS_DOM)
        r.clear_dom()
        r.raster(g, dom1, N.PASS_DOM)
        r.raster(g, dom2, N.PASS_DOM)
        out["dom_direct"] = r.dom_map().copy()
        r.dom(g, dom1, dom2)
        out["dom"] = r.dom_map()
        out["dom_status"] = r.bin_status(1)
        r.shade(cam, 1.0)
        out["shaded"] = r.shaded()
        r.filter(cam)
        out["filtered"] = r.filtered()
    finally:
        r.set_tuning(**keep)
        g.free()
    yield out


def test_heavy_lod_and_scan(heavy):
    sc = heavy["scene"]
    ref = oracle.Frame(sc.mesh, sc.w.width, sc.w.height, sc.w.dom_size)
    cap = max(1, sc.mesh.strand_count)
    lod = ref.lod(sc.params[0], sc.w.control_points, heavy["frac"])
    assert np.array_equal(heavy["cam_lod"], lod)
    refs, total = ref.scan(sc.params[0], lod, cap)
    got, gtotal = heavy["cam_refs"]
    assert gtotal == total and np.array_equal(got, refs)
    if heavy["frac"] > 0.0:  # n = ceil(fraction * (N_b + jitter)) strands per bundle
        assert abs(total / cap - heavy["frac"]) < 0.02, (total, cap, heavy["frac"])
    dlod = ref.lod(sc.params[1])
    assert np.array_equal(heavy["dom_lod"], dlod)
    drefs, dtotal = ref.scan(sc.params[1], dlod, cap)
    got, gtotal = heavy["dom_refs"]
    assert gtotal == dtotal and np.array_equal(got, drefs)


def test_heavy_gbuffer_bit_exact_against_oracle(heavy):
    sc = heavy["scene"]
    w = sc.w
    ref = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    ref.clear()
    ref.raster(sc.params[0], heavy["cam_refs"][0], threads=THREADS)
    center, cons, beta = heavy["gb"]
    assert heavy["cam_status"]["binned"] and not heavy["cam_status"]["fallback"], heavy["cam_status"]
    assert (center != EMPTY64).sum() > 10000, "nothing rasterized"
    assert np.array_equal(center, ref.center), f"{(center != ref.center).sum()} centre words differ"
    assert np.array_equal(cons, ref.cons), f"{(cons != ref.cons).sum()} conservative words differ"
    assert np.array_equal(beta, ref.beta), f"{(beta != ref.beta).sum()} beta words differ"


def test_heavy_dom_bit_exact_against_oracle(heavy):
    sc = heavy["scene"]
    w = sc.w
    ref = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    ref.clear_dom()
    refs = heavy["dom_refs"][0]
    ref.raster(sc.params[1], refs, threads=THREADS)
    ref.raster(sc.params[2], refs, threads=THREADS)
    assert heavy["dom_status"]["binned"] and not heavy["dom_status"]["fallback"], heavy["dom_status"]
    assert np.array_equal(heavy["dom"], ref.dom), f"{(heavy['dom'] != ref.dom).sum()} DOM words differ"


def test_heavy_binned_equals_direct(heavy):
    center, cons, beta = heavy["gb"]
    dc, dcons, db = heavy["gb_direct"]
    assert np.array_equal(center, dc), f"{(center != dc).sum()} centre words differ"
    assert np.array_equal(cons, dcons), f"{(cons != dcons).sum()} conservative words differ"
    assert np.array_equal(beta, db), f"{(beta != db).sum()} beta words differ"
    assert np.array_equal(heavy["dom"], heavy["dom_direct"]), \
        f"{(heavy['dom'] != heavy['dom_direct']).sum()} DOM words differ"


def test_heavy_dom_counts_total(heavy):
    S = heavy["scene"].w.dom_size
    total = int(heavy["dom"][S * S:].astype(np.uint64).sum())
    assert total == int(heavy["dom_counts1"][2]) > 0


def test_heavy_shade_and_filter_against_oracle(heavy):
    sc = heavy["scene"]
    w = sc.w
    ref = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    ref.center[:], ref.cons[:], ref.beta[:] = heavy["gb"]
    ref.dom[:] = heavy["dom"]
    pa, n_p, da, n_d = sc.light_arrays()
    ref.shade(sc.params[0], pa, n_p, da, n_d, threads=THREADS)
    shaded, depth = heavy["shaded"]
    assert np.array_equal(depth, ref.shaded_depth)
    e = rel_err(shaded, ref.shaded)
    print(f"{w.name} fraction {heavy['frac']}: shade max rel err {e.max():.3e} over "
          f"{int((depth > 0).sum())} pixels")
    assert e.max() <= RGB_TOL
    # the filter from the GPU's own shading
    ref.shaded[:], ref.shaded_depth[:] = shaded, depth
    ref.filter(sc.params[0], threads=THREADS)
    filtered, fdepth = heavy["filtered"]
    assert np.array_equal(fdepth, ref.filtered_depth)
    assert np.array_equal(filtered[:, 3], ref.filtered[:, 3]), "filter alpha differs"
    ef = rel_err(filtered[:, :3], ref.filtered[:, :3])
    print(f"{w.name} fraction {heavy['frac']}: filter max rel err {ef.max():.3e}")
    assert ef.max() <= RGB_TOL
