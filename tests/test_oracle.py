"""Pins the oracle (CPU restatement) before it is trusted as the parity target.

Sources of truth, none of which is a retyped copy of the oracle's formula:
  - the reference's own known-answer tests (src/gbuffer.rs:65-104,
    src/lod.rs:76-134) and the PCG values they imply;
  - an independent float64 restatement of src/mesh.rs::styled_position, which
    the reference documents as "matching the rasterizer" (mesh.rs:98-110);
  - hand-enumerated pixel sets for straight strands on an affine camera;
  - invariants (order independence, DOM count totals, empty sentinels);
  - textbook forms and quadrature for the shading integrals.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from bevy_hair_amd.groom import synthetic_groom
from tests.mesh_ref import styled_position
from tests.oracle_scenes import line_mesh, refs_for, screen_params

L = oracle.lib()
EMPTY64 = 0xFFFFFFFFFFFFFFFF


def u32s(n):
    return (C.c_uint32 * n)()


# ---- G-buffer word (gbuffer.rs tests) ----------------------------------------------


def test_gbuffer_roundtrip_known_word():
    # gbuffer.rs:65-76: depth 0x123456, oct [3, 250], uvw [1, 17, 63], ao 40.
    word = L.or_pack_gb(0x123456, 3, 250, 1, 17, 63, 40)
    assert word == 0x12345603FA051FE8
    out = u32s(7)
    L.or_unpack_gb(word, out)
    assert list(out) == [0x123456, 3, 250, 1, 17, 63, 40]
    assert not L.or_is_empty(word)
    assert L.or_is_empty(EMPTY64)


def test_gbuffer_nearer_is_smaller():
    # gbuffer.rs:78-94
    far = L.or_pack_gb(1000, 255, 255, 63, 63, 63, 63)
    near = L.or_pack_gb(999, 0, 0, 0, 0, 0, 0)
    assert near < far < EMPTY64


def test_quantize_depth_closed_forms():
    assert L.or_quantize_depth(0.5, 0.5, 2.5) == 0
    assert L.or_quantize_depth(2.5, 0.5, 2.5) == 0xFFFFFE
    assert L.or_quantize_depth(9.0, 0.5, 2.5) == 0xFFFFFE  # clamped below the empty sentinel
    assert L.or_quantize_depth(-3.0, 0.5, 2.5) == 0
    assert L.or_quantize_depth(1.0, 0.0, 2.0) == 8388607  # floor(0.5 * 16777214)
    assert L.or_dequantize_depth(0, 0.5, 2.5) == 0.5
    assert L.or_dequantize_depth(0xFFFFFE, 0.5, 2.5) == 2.5


# ---- PCG ---------------------------------------------------------------------------------


def test_pcg_known_values():
    assert [L.or_pcg(i) for i in range(4)] == [0x07BB2FE2, 0xA8BEEA3C, 0x7A7ECC88, 0x7F0EF6BC]


def test_pcg_uv_in_unit_square():
    out = (C.c_float * 2)()
    vals = []
    for s in range(2000):
        L.or_pcg_uv(s, out)
        vals.append(tuple(out))
    v = np.array(vals)
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert abs(v.mean() - 0.5) < 0.02  # uniform


# ---- octahedral tangent -----------------------------------------------------------------


@pytest.mark.parametrize("n,expect", [((0, 0, 1), (128, 128)), ((1, 0, 0), (255, 128)), ((-1, 0, 0), (0, 128)),
                                       ((0, 1, 0), (128, 255)), ((0, -1, 0), (128, 0)), ((0, 0, -1), (255, 255))])
def test_oct_encode_axes(n, expect):
    out = u32s(2)
    L.or_oct_encode((C.c_float * 3)(*n), out)
    assert tuple(out) == expect


def test_oct_roundtrip_angle():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(3000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    enc, dec = u32s(2), (C.c_float * 3)()
    worst = 0.0
    for n in v:
        L.or_oct_encode((C.c_float * 3)(*n), enc)
        L.or_oct_decode(enc[0], enc[1], dec)
        worst = max(worst, np.degrees(np.arccos(np.clip(np.dot(n, list(dec)), -1, 1))))
    # 8+8-bit octahedral quantization; the reference's u32(e * 255 + 0.5) has no
    # code for e = 0.5 (it decodes 128 as +0.004), so the bound is ~1 degree.
    assert worst < 1.0


# ---- LOD (lod.rs tests and Eqs. 1, 4-5) ------------------------------------------------------


def test_control_points_known_answers():
    f = L.or_control_points_for_lod
    assert f(0.0) == 9 and f(0.1) == 33 and f(0.262) == 65 and f(1.0) == 65
    for l in np.linspace(0, 1, 401):
        c = f(float(l))
        # lod.rs:92-106: C - 1 is a power of two, 3 <= C <= 127.
        assert 3 <= c <= 127 and bin(c - 1).count("1") == 1
        # ... and it is the largest 2^k + 1 not above C_raw = max(floor(sqrt(L) * 127), 16).
        c_raw = max(int(np.floor(np.sqrt(np.float32(l)) * np.float32(127.0))), 16)
        best = max(2**k + 1 for k in range(8) if 2**k + 1 <= c_raw)
        assert c == best


def _lod_for_span(span_px, n, lod=1, mode=0, offset=(100.0, 100.0), height=1080.0, behind=False):
    """LOD record of one bundle whose cage spans `span_px` pixels horizontally."""
    x0, y0 = offset
    z = -5.0 if behind else 0.0
    mesh = line_mesh((x0, y0, z), (x0 + span_px, y0, z), strands=n)
    p = screen_params(1920, int(height), mode=mode, lod=lod)
    if behind:
        p.camera_forward[:] = [0.0, 0.0, 1.0, 0.0]
    f = oracle.Frame(mesh, 1920, int(height), 64)
    return f.lod(p)[0], p


def _delta(bundle_id=0):
    return L.or_pcg(bundle_id ^ 0x9E3779B9) * np.float32(1.0 / 4294967296.0)


def test_lod_selection_and_strand_count():
    # lod.rs:76-90. L = diag / R_y * lambda (lambda = 3 in screen_params).
    rec, _ = _lod_for_span(1080.0, 100)  # L = 3 -> clamped to 1
    assert rec[0] == 100 and rec[1] == 65 and rec[2] == 100
    rec, _ = _lod_for_span(100.0, 100)  # L = 300 / 1080
    l = np.float32(100.0) / np.float32(1080.0) * np.float32(3.0)
    assert abs(float(l) - 300.0 / 1080.0) < 1e-6
    assert rec[0] == int(np.ceil(l * (np.float32(100.0) + _delta())))
    assert rec[1] == L.or_control_points_for_lod(float(l))
    rec, _ = _lod_for_span(0.0, 100)  # L = 0: at least one strand survives
    assert rec[0] == 1 and rec[1] == 9


def test_lod_cull_dom_and_lod_off():
    rec, _ = _lod_for_span(50.0, 40, offset=(-500.0, 100.0))  # AABB left of the viewport
    assert rec[0] == 0 and rec[1] == 16 and rec[2] == 40
    rec, _ = _lod_for_span(50.0, 40, behind=True)  # behind the camera
    assert rec[0] == 0
    rec, _ = _lod_for_span(5.0, 40, mode=1)  # DOM pass: full strand set at C(1)
    assert rec[0] == 40 and rec[1] == 65
    rec, _ = _lod_for_span(5.0, 40, lod=0)  # LOD off: full set at 127 points
    assert rec[0] == 40 and rec[1] == 127


def test_scan_order_and_capacity():
    p = screen_params(64, 64, bundles=3)
    lod = np.array([[2, 9, 5, 0], [0, 16, 4, 0], [3, 17, 3, 0]], np.uint32)
    refs, total = oracle.Frame.scan(p, lod, 100)
    assert total == 5
    assert refs.tolist() == [[0, 0, 2, 9], [0, 1, 2, 9], [2, 0, 3, 17], [2, 1, 3, 17], [2, 2, 3, 17]]
    refs, total = oracle.Frame.scan(p, lod, 3)  # scan.wgsl:14-16 caps at the buffer length
    assert total == 3 and refs[-1].tolist() == [2, 0, 3, 17]


def test_dom_optical_depth_known_answers():
    # lod.rs:123-134
    layers = (C.c_uint32 * 16)(100, 40, 7)
    f = L.or_dom_optical_depth
    assert f(layers, 0.0) == 0.0 and f(layers, -1.0) == 0.0
    assert abs(f(layers, 0.5) - 50.0) < 1e-4
    assert abs(f(layers, 1.0) - 100.0) < 1e-4
    assert abs(f(layers, 1.25) - 110.0) < 1e-4
    assert abs(f(layers, 3.0) - 147.0) < 1e-4
    assert abs(f(layers, 3.0) - 147.0) < 1e-4


# ---- strand generation vs mesh.rs::styled_position -----------------------------------------


def test_eval_cp_matches_mesh_rs_mirror():
    mesh = synthetic_groom("wavy", 2048, 11)
    rng = np.random.default_rng(3)
    out = (C.c_float * 4)()
    worst = 0.0
    for _ in range(300):
        b = int(rng.integers(mesh.bundle_count))
        u, v = rng.uniform(0, 1, 2)
        ncp = int(rng.choice([9, 17, 33, 65, 127]))
        idx = int(rng.integers(ncp))
        L.or_eval_cp(mesh.bundles.ctypes.data, mesh.layers.ctypes.data, mesh.style.ctypes.data, b,
                     float(np.float32(u)), float(np.float32(v)), idx, ncp, out)
        ref = styled_position(mesh, b, (float(np.float32(u)), float(np.float32(v))), idx / (ncp - 1))
        worst = max(worst, float(np.abs(np.array(list(out)[:3]) - ref).max()))
    assert worst < 2e-5


# ---- raster: hand-enumerated pixel sets -------------------------------------------------------


def raster_line(a, b, ncp, mode=0, width=128, height=96, scene=None, reverse=1, n_lod=None, strands=1,
                center_diameter=1.0):
    mesh = line_mesh(a, b, strands=strands)
    f = oracle.Frame(mesh, width, height, width)
    if scene is not None:
        f.scene_depth = np.ascontiguousarray(scene, np.float32)
    f.clear()
    f.clear_dom()
    refs = refs_for(mesh, ncp, n_lod)
    if mode == 0:
        counts = f.raster(screen_params(width, height, reverse=reverse, center_diameter=center_diameter), refs)
    else:
        c1 = f.raster(screen_params(width, height, mode=1), refs)
        counts = f.raster(screen_params(width, height, mode=2), refs)
        assert c1[2] == counts[2]
    return f, counts


def pixels(buf, width):
    idx = np.nonzero(buf != np.uint64(EMPTY64))[0] if buf.dtype == np.uint64 else np.nonzero(buf)[0]
    return {(int(i % width), int(i // width)) for i in idx}


def test_raster_horizontal_line_pixel_set():
    f, counts = raster_line((2.5, 3.5, 0.0), (66.5, 3.5, 0.0), ncp=17)
    expect = {(x, 3) for x in range(2, 67)}
    assert pixels(f.center, 128) == expect
    assert pixels(f.cons, 128) == expect
    # every sample is a centre hit; 16 segments of ~4 px give 5 (or, when rounding
    # lands just past 4.0 px, 6) samples each
    assert counts[0] == counts[1] and 16 * 5 <= counts[0] <= 16 * 6 and counts[3] == 16
    word = int(f.center[3 * 128 + 10])
    assert (word >> 32) & 0xFF == 255 and (word >> 24) & 0xFF == 128  # tangent +x -> oct (255, 128)
    depth = word >> 40
    assert depth == L.or_quantize_depth(1.0, 0.05, 10.0)
    # beta word = depth << 8 | round(n_lod / N * 255) with n_lod = N
    assert int(f.beta[3 * 128 + 10]) == (depth << 8) | 255
    assert pixels(f.beta, 128) == expect or all(f.beta[y * 128 + x] != 0xFFFFFFFF for x, y in expect)


def test_raster_diagonal_adds_four_connected_conservative_pixels():
    f, _ = raster_line((2.5, 2.5, 0.0), (34.5, 34.5, 0.0), ncp=17)
    diag = {(k, k) for k in range(2, 35)}
    assert pixels(f.center, 128) == diag
    extra = {(k - 1, k) for k in range(3, 35)} | {(k, k - 1) for k in range(3, 35)}
    assert pixels(f.cons, 128) == diag | extra


def test_raster_beta_word_half_lod():
    f, _ = raster_line((2.5, 3.5, 0.0), (18.5, 3.5, 0.0), ncp=9, strands=4, n_lod=2)
    word = int(f.beta[3 * 128 + 5])
    assert word & 0xFF == 128  # u32(0.5 * 255 + 0.5)


def test_raster_rejects_near_offscreen_and_occluded():
    # behind the near plane: view distance 0.03 < near 0.05
    f, _ = raster_line((2.5, 3.5, -0.97), (30.5, 3.5, -0.97), ncp=17)
    assert not pixels(f.cons, 128)
    # entirely off screen
    f, _ = raster_line((-80.5, 3.5, 0.0), (-10.5, 3.5, 0.0), ncp=17)
    assert not pixels(f.cons, 128)
    # reverse-Z scene depth 0.6 over x < 30 occludes clip depth 0.5 there
    scene = np.zeros((96, 128), np.float32)
    scene[:, :30] = 0.6
    f, _ = raster_line((2.5, 3.5, 0.0), (66.5, 3.5, 0.0), ncp=17, scene=scene)
    assert pixels(f.cons, 128) == {(x, 3) for x in range(30, 67)}
    # forward-Z: the same depth now passes where scene >= 0.5
    f, _ = raster_line((2.5, 3.5, 0.0), (66.5, 3.5, 0.0), ncp=17, scene=scene, reverse=0)
    assert pixels(f.cons, 128) == {(x, 3) for x in range(2, 30)}


def test_raster_center_diameter_gates_center_hits():
    # a line through pixel corners: centres are 0.5 px away -> only conservative with a thin disk
    f, _ = raster_line((2.5, 3.0, 0.0), (66.5, 3.0, 0.0), ncp=17, center_diameter=0.5)
    assert not pixels(f.center, 128) and pixels(f.cons, 128)
    f, _ = raster_line((2.5, 3.0, 0.0), (66.5, 3.0, 0.0), ncp=17, center_diameter=1.0)
    assert pixels(f.center, 128)


def test_dom_bresenham_counts_each_pixel_once():
    f, counts = raster_line((2.5, 3.5, 0.0), (66.5, 3.5, 0.0), ncp=17, mode=1, width=128, height=128)
    S = 128
    front = f.dom[: S * S]
    expect = {(x, 3) for x in range(2, 67)}
    assert {(int(i % S), int(i // S)) for i in np.nonzero(front != 0xFFFFFFFF)[0]} == expect
    cnt = f.dom[S * S:].reshape(S * S, 16)
    # shared segment end points are binned once; all hair sits in the front slice
    assert counts[2] == 65 and cnt.sum() == 65
    assert all(cnt[y * S + x, 0] == 1 for x, y in expect)


# ---- invariants on a real groom ------------------------------------------------------------------


@pytest.fixture(scope="module")
def c1_scene():
    from tests.scenes import make_scene
    return make_scene("c1")


def test_raster_order_independent_and_threadsafe(c1_scene):
    sc = c1_scene
    w = sc.w
    f = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    lod = f.lod(sc.params[0], w.control_points, w.strand_fraction)
    refs, _ = f.scan(sc.params[0], lod, sc.mesh.strand_count)
    f.clear()
    f.raster(sc.params[0], refs)
    a = (f.center.copy(), f.cons.copy(), f.beta.copy())
    f.clear()
    f.raster(sc.params[0], refs[np.random.default_rng(1).permutation(len(refs))].copy())
    assert all(np.array_equal(x, y) for x, y in zip(a, (f.center, f.cons, f.beta)))
    f.clear()
    f.raster(sc.params[0], refs, threads=4)
    assert all(np.array_equal(x, y) for x, y in zip(a, (f.center, f.cons, f.beta)))


def test_dom_totals_equal_depth_pass_writes(c1_scene):
    sc = c1_scene
    w = sc.w
    f = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    lod = f.lod(sc.params[1])
    refs, _ = f.scan(sc.params[1], lod, sc.mesh.strand_count)
    f.clear_dom()
    c1 = f.raster(sc.params[1], refs)
    c2 = f.raster(sc.params[2], refs)
    S = w.dom_size
    assert c1[2] == c2[2] == f.dom[S * S:].sum() > 0
    front = f.dom[: S * S]
    counts = f.dom[S * S:].reshape(-1, 16)
    # texels without a front depth hold no counts; texels with one hold >= 1 in slice 0
    assert counts[front == 0xFFFFFFFF].sum() == 0
    assert (counts[front != 0xFFFFFFFF, 0] >= 1).all()


# ---- shading: textbook forms and quadrature ------------------------------------------------------


def test_fresnel_schlick_ior_155():
    r0 = ((1 - 1.55) / (1 + 1.55)) ** 2
    assert abs(L.or_fresnel(1.0) - r0) < 1e-7
    assert abs(L.or_fresnel(0.0) - 1.0) < 1e-7
    assert abs(L.or_fresnel(0.5) - (r0 + (1 - r0) * 0.5**5)) < 1e-7


def test_absorption_inverts_chiang_eq9():
    # Chiang et al. 2016 Eq. 9: sigma_a = (ln c / (5.969 - 0.215 b + 2.532 b^2 - 10.73 b^3 + 5.574 b^4 + 0.245 b^5))^2
    out = (C.c_double * 3)()
    for b in (0.0, 0.2, 0.5, 1.0):
        c = (0.9, 0.3, 0.05)
        L.or_absorption_from_color((C.c_double * 3)(*c), b, out)
        d = 5.969 - 0.215 * b + 2.532 * b**2 - 10.73 * b**3 + 5.574 * b**4 + 0.245 * b**5
        # the shader's f32 literals round the coefficients
        assert np.allclose(list(out), [(np.log(x) / d) ** 2 for x in c], rtol=1e-6)
    L.or_absorption_from_color((C.c_double * 3)(1.0, 1.0, 1.0), 0.3, out)
    assert list(out) == [0.0, 0.0, 0.0]


def az(ctd, cphi, sigma):
    out = (C.c_double * 7)()
    L.or_azimuthal(ctd, cphi, (C.c_double * 3)(*sigma), out)
    return np.array(list(out))


def test_azimuthal_karis_forms():
    sigma = (0.3, 0.6, 1.2)
    ctd = 0.9
    # R: 0.25 cos(phi/2) F(cos(theta_d) cos(phi/2)); at phi = 0 it is 0.25 F(cos theta_d)
    assert abs(az(ctd, 1.0, sigma)[0] - 0.25 * L.or_fresnel(ctd)) < 1e-12
    # TT peaks backward (cos phi = -1), TRT forward (cos phi = 1)
    assert az(ctd, -1.0, sigma)[1] > 100 * az(ctd, 1.0, sigma)[1]
    assert az(ctd, 1.0, sigma)[4] > 100 * az(ctd, -1.0, sigma)[4]
    # absorption darkens TT/TRT per channel in sigma order
    v = az(ctd, 0.0, sigma)
    assert v[1] > v[2] > v[3] and v[4] > v[5] > v[6]


def test_dual_scatter_midpoint_matches_quadrature():
    from scipy.integrate import quad

    sigma = (0.2, 0.5, 1.5)
    ctd = 0.8
    out = (C.c_double * 24)()
    L.or_dual_scatter(ctd, (C.c_double * 3)(*sigma), 0.3, 0.08, out)
    ds = np.array(list(out)).reshape(8, 3)
    # fwd lobe energy = 2 * integral over the forward quarter turn phi in [pi/2, pi]
    for lobe, col in ((0, 0), (1, 1), (2, 4)):
        for ch in range(3):
            k = col + (ch if lobe else 0)
            val, _ = quad(lambda phi: az(ctd, np.cos(phi), sigma)[k], np.pi / 2, np.pi, epsabs=1e-12)
            assert abs(ds[lobe, ch] - 2 * val) <= 2e-2 * abs(2 * val) + 1e-9
    # a_f = min(sum of forward lobes, 0.95)
    assert np.allclose(ds[3], np.minimum(ds[0] + ds[1] + ds[2], 0.95))


def test_bsdf_depth_limits():
    t = (0.0, 1.0, 0.0)
    v = (0.0, 0.2, 0.98)
    l = (0.5, 0.3, 0.81)
    alb = (0.3, 0.15, 0.08)
    out = (C.c_double * 3)()

    def bsdf(depth):
        L.or_hair_bsdf((C.c_double * 3)(*t), (C.c_double * 3)(*v), (C.c_double * 3)(*l), (C.c_double * 3)(*alb),
                       0.3, 0.08, depth, out)
        return np.array(list(out))

    d0 = bsdf(0.0)
    assert (d0 > 0).all()
    # below the 1e-3 threshold there is no global term: exp(-depth) scales the local terms only
    assert np.allclose(bsdf(5e-4), d0 * np.exp(-5e-4), rtol=1e-12)
    # very deep inside the groom nothing arrives
    assert (bsdf(400.0) < 1e-6 * d0).all()


# ---- filter ------------------------------------------------------------------------------------------


def filter_frame(center_mask, cons_mask, color, width=32, height=24, depth=1000):
    mesh = line_mesh((0, 0, 0), (1, 0, 0))
    f = oracle.Frame(mesh, width, height, width)
    word = L.or_pack_gb(depth, 255, 128, 0, 0, 0, 63)  # tangent +x on screen
    f.center[:] = np.where(center_mask.reshape(-1), np.uint64(word), np.uint64(EMPTY64))
    f.cons[:] = np.where((center_mask | cons_mask).reshape(-1), np.uint64(word), np.uint64(EMPTY64))
    f.shaded[:] = 0
    f.shaded[center_mask.reshape(-1)] = [*color, 1.0]
    f.filter(screen_params(width, height))
    return f


def test_filter_preserves_constant_colour_and_full_coverage():
    m = np.ones((24, 32), bool)
    f = filter_frame(m, np.zeros_like(m), (0.25, 0.5, 0.125))
    assert np.allclose(f.filtered[:, :3], [0.25, 0.5, 0.125], rtol=1e-6, atol=0)
    assert (f.filtered[:, 3] == 1.0).all()


def test_filter_coverage_is_center_over_conservative_in_strip():
    m = np.zeros((24, 32), bool)
    c = np.zeros_like(m)
    m[10, 5:11] = True    # 6 centre hits left of and at (10, 10)
    c[10, 11:16] = True   # 5 conservative-only hits to the right
    f = filter_frame(m, c, (0.2, 0.2, 0.2))
    alpha = f.filtered[10 * 32 + 10, 3]
    assert abs(alpha - 6.0 / 11.0) < 1e-7
    # an empty pixel stays empty
    assert (f.filtered[0] == 0).all()


def test_fp32_against_fp64_error_statement(capsys):
    """The float restatement (what the GPU path must match) against its float64
    instantiation, stage by stage on the same inputs: the stated error of each
    output of the path."""
    from tests.scenes import make_scene, rel_err
    sc = make_scene("small_lod")
    w = sc.w
    f32 = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    f64 = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    cap = max(1, sc.mesh.strand_count)
    cam, dom1, dom2 = sc.params
    lod = f32.lod(cam, w.control_points, w.strand_fraction)
    lod64 = f64.lod(cam, w.control_points, w.strand_fraction, f64=True)
    refs, _ = f32.scan(cam, lod, cap)
    f32.clear(); f64.clear()
    f32.raster(cam, refs)
    f64.raster(cam, refs, f64=True)
    depth_step = np.abs((f32.center >> np.uint64(40)).astype(np.int64) - (f64.center >> np.uint64(40)).astype(np.int64))
    both = (f32.center != np.uint64(EMPTY64)) & (f64.center != np.uint64(EMPTY64))
    gb_diff = float((f32.center != f64.center).mean())
    dlod = f32.lod(dom1)
    drefs, _ = f32.scan(dom1, dlod, cap)
    f32.clear_dom(); f64.clear_dom()
    for p in (dom1, dom2):
        f32.raster(p, drefs)
        f64.raster(p, drefs, f64=True)
    dom_diff = float((f32.dom != f64.dom).mean())
    # shading and filtering on identical (float-path) inputs
    f64.center[:], f64.cons[:], f64.beta[:], f64.dom[:] = f32.center, f32.cons, f32.beta, f32.dom
    pa, n_p, da, n_d = sc.light_arrays()
    f32.shade(cam, pa, n_p, da, n_d)
    f64.shade(cam, pa, n_p, da, n_d, f64=True)
    e_shade = rel_err(f32.shaded, f64.shaded)
    f64.shaded[:], f64.shaded_depth[:] = f32.shaded, f32.shaded_depth
    f32.filter(cam)
    f64.filter(cam, f64=True)
    e_filter = rel_err(f32.filtered, f64.filtered)
    with capsys.disabled():
        print(f"\n[fp32 vs fp64] lod records equal: {np.array_equal(lod, lod64)}; "
              f"G-buffer words differing: {gb_diff:.2e} (max depth difference {int(depth_step[both].max()) if both.any() else 0} "
              f"quantization steps); DOM words differing: {dom_diff:.2e}; shaded RGB max rel {e_shade.max():.2e}; "
              f"filtered max rel {e_filter.max():.2e}")
    # sanity bounds of the statement (small_lod, seeded): float rounding moves a
    # few percent of the words' low bits and the depth key by at most a couple
    # of quantization steps
    assert np.array_equal(lod, lod64)
    assert gb_diff < 0.1 and (not both.any() or depth_step[both].max() <= 4) and dom_diff < 0.02
    assert e_shade.max() < 1e-2 and e_filter.max() < 1e-3
