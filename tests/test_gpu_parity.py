"""HIP path vs the oracle on identical seeded inputs (bit-exact keys, 1e-3 RGB).

Each stage is checked in isolation: the oracle stage consumes the GPU's own
outputs of the previous stage, so a failure names the kernel that diverged.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from tests.scenes import EMPTY64, make_scene, rel_err

pytestmark = pytest.mark.gpu

CASES = ["tiny", "c1", "small_lod"]
RGB_TOL = 1e-3


@pytest.fixture(scope="module", params=CASES)
def run(request, renderer):
    sc = make_scene(request.param)
    w = sc.w
    print(f"\n[gpu parity] case={request.param} strands={w.strands} {w.width}x{w.height} dom={w.dom_size}", flush=True)
    r = renderer
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    cam, dom1, dom2 = sc.params[0], sc.params[1], sc.params[2]
    out = {"scene": sc, "g": g}
    r.clear()
    r.lod(g, cam, N.PASS_CAMERA, w.control_points, w.strand_fraction)
    r.scan(g, cam, N.PASS_CAMERA)
    out["cam_lod"] = r.lod_records(g, N.PASS_CAMERA)
    out["cam_refs"] = r.refs(g, N.PASS_CAMERA)
    r.raster(g, cam, N.PASS_CAMERA)
    out["gb"] = r.gbuffer()
    out["cam_counts"] = r.raster_count(g, cam, N.PASS_CAMERA)
    keep = r.get_tuning()
    r.set_tuning(dom_tile=0)  # kernel-by-kernel DOM below; hm_dom is tested separately
    r.clear_dom()
    r.lod(g, dom1, N.PASS_DOM)
    r.scan(g, dom1, N.PASS_DOM)
    out["dom_lod"] = r.lod_records(g, N.PASS_DOM)
    out["dom_refs"] = r.refs(g, N.PASS_DOM)
    out["dom_counts1"] = r.raster_count(g, dom1, N.PASS_DOM)
    r.raster(g, dom1, N.PASS_DOM)
    out["dom_front"] = r.dom_map()
    out["dom_counts2"] = r.raster_count(g, dom2, N.PASS_DOM)
    r.raster(g, dom2, N.PASS_DOM)
    out["dom"] = r.dom_map()
    r.shade(cam, 1.0)
    out["shaded"] = r.shaded()
    r.filter(cam)
    out["filtered"] = r.filtered()
    r.set_tuning(**keep)
    yield out
    g.free()


def ref_frame(sc):
    return oracle.Frame(sc.mesh, sc.w.width, sc.w.height, sc.w.dom_size)


def test_lod_and_scan(run):
    sc = run["scene"]
    ref = ref_frame(sc)
    cap = max(1, sc.mesh.strand_count)
    lod = ref.lod(sc.params[0], sc.w.control_points, sc.w.strand_fraction)
    assert np.array_equal(run["cam_lod"], lod)
    refs, total = ref.scan(sc.params[0], lod, cap)
    got, gtotal = run["cam_refs"]
    assert gtotal == total and np.array_equal(got, refs)
    dlod = ref.lod(sc.params[1])
    assert np.array_equal(run["dom_lod"], dlod)
    drefs, dtotal = ref.scan(sc.params[1], dlod, cap)
    got, gtotal = run["dom_refs"]
    assert gtotal == dtotal and np.array_equal(got, drefs)


def test_camera_gbuffer_bit_exact(run):
    sc = run["scene"]
    ref = ref_frame(sc)
    ref.clear()
    counts = ref.raster(sc.params[0], run["cam_refs"][0])
    center, cons, beta = run["gb"]
    assert (center != EMPTY64).sum() > 0, "nothing rasterized"
    assert np.array_equal(center, ref.center), f"{(center != ref.center).sum()} centre words differ"
    assert np.array_equal(cons, ref.cons), f"{(cons != ref.cons).sum()} conservative words differ"
    assert np.array_equal(beta, ref.beta), f"{(beta != ref.beta).sum()} beta words differ"
    assert np.array_equal(run["cam_counts"][:2], counts[:2])
    assert run["cam_counts"][3] == counts[3]


def test_dom_bit_exact(run):
    sc = run["scene"]
    ref = ref_frame(sc)
    S = sc.w.dom_size
    ref.clear_dom()
    c1 = ref.raster(sc.params[1], run["dom_refs"][0])
    assert np.array_equal(run["dom_front"][: S * S], ref.dom[: S * S])
    c2 = ref.raster(sc.params[2], run["dom_refs"][0])
    assert np.array_equal(run["dom"], ref.dom), f"{(run['dom'] != ref.dom).sum()} DOM words differ"
    assert run["dom_counts1"][2] == c1[2] and run["dom_counts2"][2] == c2[2]
    # every depth-pass write is one opacity increment
    assert int(ref.dom[S * S:].sum()) == int(c2[2]) == int(c1[2])


def test_shade_within_tolerance(run):
    sc = run["scene"]
    ref = ref_frame(sc)
    ref.center[:], ref.cons[:], ref.beta[:] = run["gb"]
    ref.dom[:] = run["dom"]
    pa, n_p, da, n_d = sc.light_arrays()
    ref.shade(sc.params[0], pa, n_p, da, n_d)
    shaded, depth = run["shaded"]
    assert np.array_equal(depth, ref.shaded_depth)
    e = rel_err(shaded, ref.shaded)
    print(f"shade max rel err {e.max():.3e}")
    assert e.max() <= RGB_TOL


def test_filter_within_tolerance(run):
    sc = run["scene"]
    ref = ref_frame(sc)
    ref.center[:], ref.cons[:], ref.beta[:] = run["gb"]
    ref.shaded[:], ref.shaded_depth[:] = run["shaded"]
    ref.filter(sc.params[0])
    filt, depth = run["filtered"]
    assert np.array_equal(depth, ref.filtered_depth)
    assert np.array_equal(filt[:, 3], ref.filtered[:, 3]), "coverage alpha differs"
    e = rel_err(filt, ref.filtered)
    print(f"filter max rel err {e.max():.3e}")
    assert e.max() <= RGB_TOL


def test_frame_matches_oracle_and_is_deterministic(renderer):
    sc = make_scene("c1")
    w, r = sc.w, renderer
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    opts = r.frame_opts(control_points=w.control_points)
    r.hair_pass(g, sc.params, opts)
    first = (*r.gbuffer(), r.dom_map(), r.filtered()[0])
    r.hair_pass(g, sc.params, opts)
    second = (*r.gbuffer(), r.dom_map(), r.filtered()[0])
    g.free()
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
    ref = ref_frame(sc)
    pa, n_p, da, n_d = sc.light_arrays()
    ref.hair_pass(sc.params, pa, n_p, da, n_d, control_points=w.control_points)
    assert np.array_equal(first[0], ref.center)
    assert np.array_equal(first[1], ref.cons)
    assert np.array_equal(first[2], ref.beta)
    assert np.array_equal(first[3], ref.dom)
    assert rel_err(first[4], ref.filtered).max() <= RGB_TOL


@pytest.mark.parametrize("cam_tile,chunk", [(16, 1024), (32, 64), (32, 2048)])
def test_frame_with_long_camera_segments(renderer, cam_tile, chunk):
    """A box limit of one tile sends most segments down the long path, which
    draws after the tiles have emitted the pixel lists: pixels only long
    segments cover must still be shaded and filtered like the oracle's."""
    sc = make_scene("c1")
    w, r = sc.w, renderer
    print(f"\n[gpu parity] frame with long segments cam_tile={cam_tile} chunk={chunk}", flush=True)
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    keep = r.get_tuning()
    try:
        r.set_tuning(cam_tile=cam_tile, bin_box_limit=1, cam_chunk=chunk)
        r.hair_pass(g, sc.params, r.frame_opts(control_points=w.control_points))
        st = r.bin_status()
        gb = r.gbuffer()
        shaded, filtered = r.shaded()[0], r.filtered()[0]
    finally:
        r.set_tuning(**keep)
        g.free()
    assert st["binned"] and st["long"] and st["long_count"] > 0
    ref = ref_frame(sc)
    pa, n_p, da, n_d = sc.light_arrays()
    ref.hair_pass(sc.params, pa, n_p, da, n_d, control_points=w.control_points)
    assert np.array_equal(gb[0], ref.center) and np.array_equal(gb[1], ref.cons)
    # every covered pixel shaded and filtered, every other one cleared
    assert rel_err(shaded, ref.shaded).max() <= RGB_TOL
    assert rel_err(filtered, ref.filtered).max() <= RGB_TOL


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_raster_variants_give_identical_bytes(renderer, flags):
    """Every launch choice the autotuner may pick writes the same G-buffer and DOM."""
    sc = make_scene("small_lod")
    w, r = sc.w, renderer
    print(f"\n[gpu parity] raster flags={flags}", flush=True)
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    keep = r.get_tuning()
    try:
        for bpc in (4, 32):
            r.set_tuning([flags] * 3, bpc, dom_tile=0)
            r.hair_pass(g, sc.params, r.frame_opts(control_points=w.control_points))
            got = (*r.gbuffer(), r.dom_map())
            ref = ref_frame(sc)
            pa, n_p, da, n_d = sc.light_arrays()
            ref.hair_pass(sc.params, pa, n_p, da, n_d, control_points=w.control_points)
            for a, b in zip(got, (ref.center, ref.cons, ref.beta, ref.dom)):
                assert np.array_equal(a, b)
    finally:
        r.set_tuning(**keep)
        g.free()


_DOM_REF = {}


def oracle_dom(name):
    """The oracle's DOM after both passes over its own DOM work list."""
    if name not in _DOM_REF:
        sc = make_scene(name)
        ref = ref_frame(sc)
        cap = max(1, sc.mesh.strand_count)
        lod = ref.lod(sc.params[1])
        refs, _ = ref.scan(sc.params[1], lod, cap)
        ref.clear_dom()
        counts = ref.raster(sc.params[1], refs)
        ref.raster(sc.params[2], refs)
        _DOM_REF[name] = (ref.dom.copy(), int(counts[2]))
    return _DOM_REF[name]


# (tile, box limit, record buffer, binning workgroups per CU, persistent DOM
# tile workgroups per CU): binned with few / many long segments, short and long
# segment runs per workgroup, a record buffer too small, which must switch to
# the direct kernels on the device, and the 16-texel tiles as one workgroup
# per tile and as the persistent busy-tile kernel.
DOM_VARIANTS = [(16, 8, 0, 2, 2), (32, 8, 0, 4, 2), (16, 1, 0, 2, 2), (32, 64, 0, 1, 2), (16, 8, 0, 16, 1),
                (32, 8, 1, 2, 2), (16, 8, 0, 2, 0), (16, 1, 0, 2, 0)]


@pytest.mark.parametrize("case", ["tiny", "c1", "small_lod", "dense_dom"])
@pytest.mark.parametrize("variant", DOM_VARIANTS, ids=[f"t{t}-box{b}-rec{n}-wg{w}-busy{k}" for t, b, n, w, k in DOM_VARIANTS])
def test_binned_dom_bit_exact(renderer, case, variant):
    tile, box, records, wg, busy = variant
    sc = make_scene(case)
    w, r = sc.w, renderer
    print(f"\n[gpu parity] hm_dom case={case} tile={tile} box={box} records={records} wg={wg} busy={busy}", flush=True)
    r.targets(w.width, w.height, w.dom_size)
    g = r.upload(sc.mesh)
    keep = r.get_tuning()
    try:
        r.set_tuning(dom_tile=tile, bin_box_limit=box, bin_records=records, bin_blocks_per_cu=wg,
                     dom_busy_blocks_per_cu=busy)
        r.lod(g, sc.params[1], N.PASS_DOM)
        r.scan(g, sc.params[1], N.PASS_DOM)
        # twice: the second call must not depend on what the first left behind
        for _ in range(2):
            r.dom(g, sc.params[1], sc.params[2])
            got = r.dom_map()
            st = r.dom_bin_status()
            want, writes = oracle_dom(case)
            assert np.array_equal(got, want), f"{(got != want).sum()} DOM words differ"
        assert st["fallback"] == (records == 1) and st["binned"] == (records != 1)
        if records != 1:
            assert st["records"] > 0
        if box == 1 and records != 1:
            assert st["long_count"] > 0 and st["long"]
        S = w.dom_size
        assert int(got[S * S:].sum()) == writes
    finally:
        r.set_tuning(**keep)
        g.free()


_GB_REF = {}


def oracle_gbuffer(name, scene_depth=None):
    """The oracle's G-buffer after the camera pass over its own camera work list."""
    key = (name, scene_depth is not None)
    if key not in _GB_REF:
        sc = make_scene(name)
        ref = ref_frame(sc)
        cap = max(1, sc.mesh.strand_count)
        lod = ref.lod(sc.params[0], sc.w.control_points, sc.w.strand_fraction)
        refs, _ = ref.scan(sc.params[0], lod, cap)
        ref.clear()
        if scene_depth is not None:
            ref.scene_depth = scene_depth
        counts = ref.raster(sc.params[0], refs)
        _GB_REF[key] = (ref.center.copy(), ref.cons.copy(), ref.beta.copy(), counts)
    return _GB_REF[key]


def depth_checker(w):
    """A scene depth of 8x8 blocks that alternately keep every hair fragment
    (0: reverse-Z keeps clip >= scene) and reject every one (1: hair clip depth
    near / distance is below 1)."""
    y, x = np.mgrid[0:w.height, 0:w.width]
    return np.ascontiguousarray(np.where((x // 8 + y // 8) % 2 == 0, 0.0, 1.0).astype(np.float32).ravel())


# (tile, box limit, record buffer, binning workgroups per CU, records per tile
# workgroup, scene depth): chunk 64 splits nearly every covered tile
CAM_VARIANTS = [(16, 8, 0, 2, 1024, False), (32, 8, 0, 4, 1024, False), (16, 1, 0, 2, 64, False),
                (32, 64, 0, 1, 0, False), (16, 8, 0, 2, 64, True), (16, 8, 1, 2, 1024, False)]


@pytest.mark.parametrize("case", ["tiny", "c1", "small_lod"])
@pytest.mark.parametrize("variant", CAM_VARIANTS,
                         ids=[f"t{t}-box{b}-rec{n}-wg{w}-ch{c}{'-depth' if d else ''}"
                              for t, b, n, w, c, d in CAM_VARIANTS])
def test_binned_gbuffer_bit_exact(renderer, case, variant):
    tile, box, records, wg, chunk, use_depth = variant
    sc = make_scene(case)
    w, r = sc.w, renderer
    print(f"\n[gpu parity] hm_gbuffer case={case} tile={tile} box={box} records={records} wg={wg} "
          f"chunk={chunk} depth={use_depth}", flush=True)
    r.targets(w.width, w.height, w.dom_size)
    depth = depth_checker(w) if use_depth else None
    r.set_scene_depth(depth, 0.0)
    g = r.upload(sc.mesh)
    keep = r.get_tuning()
    try:
        r.set_tuning(cam_tile=tile, bin_box_limit=box, bin_records=records, bin_blocks_per_cu=wg, cam_chunk=chunk,
                     cam_bin_blocks_per_cu=wg)
        r.lod(g, sc.params[0], N.PASS_CAMERA, w.control_points, w.strand_fraction)
        r.scan(g, sc.params[0], N.PASS_CAMERA)
        want = oracle_gbuffer(case, depth)
        for _ in range(2):  # the second call must not depend on what the first left behind
            r.gbuffer_pass(g, sc.params[0])
            got = r.gbuffer()
            st = r.bin_status()
            for a, b, name in zip(got, want[:3], ("centre", "conservative", "beta")):
                assert np.array_equal(a, b), f"{(a != b).sum()} {name} words differ"
        assert (got[0] != EMPTY64).sum() > 0
        assert st["fallback"] == (records == 1) and st["binned"] == (records != 1)
        if box == 1 and records != 1:
            assert st["long_count"] > 0
        if use_depth:
            full = oracle_gbuffer(case)
            assert 0 < want[3][0] < full[3][0], "the depth blocks should reject some but not all writes"
    finally:
        r.set_tuning(**keep)
        r.set_scene_depth(None, 0.0)
        g.free()
