"""Pins of the oracle's environment probe, cube map sampling and rect lights
(shade.wesl:329-427, 503-518), each checked against something other than
the restatement's own formula:

- cube_face against the inverse of Vulkan's face selection table
  (scene.cube_texel_dirs builds directions from face coordinates);
- cube_sample: texel centres return the texel, the midpoint of two texels
  their mean, a clamped edge the edge texel, and levels blend linearly in
  the level (a map whose level m holds the constant m reads the level back);
- cubemap_dir against scipy's quaternion rotation, then z negated;
- probe_radiance: a grey albedo under a constant diffuse map gives
  albedo x irradiance x intensity (the tint is one); the specular weight at
  normal incidence is Schlick's f0 = ((n-1)/(n+1))^2 plus TRT; the sampled
  face follows the tilt (theta_i = tilt - theta_o); roughness selects level
  roughness x (mips - 1);
- rect lights: a small rect facing the pixel is a point light of colour
  w x h x colour; a rect facing away, or out of range, adds nothing; the
  contribution is linear in the area.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from bevy_hair_amd import plugin, scene
from bevy_hair_amd.groom import synthetic_groom

L = oracle.lib()
EMPTY64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def out3():
    return (C.c_double * 3)()


def face_of(d):
    o = out3()
    L.or_cube_face(d3(d), o)
    return int(o[0]), o[1], o[2]


def test_cube_face_inverts_the_vulkan_table():
    rng = np.random.default_rng(3)
    size = 16
    dirs = scene.cube_texel_dirs(size)
    for f in range(6):
        for _ in range(40):
            y, x = rng.integers(0, size, 2)
            face, s, t = face_of(dirs[f, y, x])
            assert face == f
            assert s == pytest.approx((x + 0.5) / size, abs=1e-12)
            assert t == pytest.approx((y + 0.5) / size, abs=1e-12)
    # the major axis decides: +x, -x, +y, -y, +z, -z
    for f, d in enumerate([(1, .2, .3), (-1, .2, .3), (.2, 1, .3), (.2, -1, .3), (.2, .3, 1), (.2, .3, -1)]):
        assert face_of(d)[0] == f


def random_cube(rng, size, mips=1):
    return [rng.uniform(0.0, 1.0, (6, max(size >> m, 1), max(size >> m, 1), 4)).astype(np.float32)
            for m in range(mips)]


def sample(levels, d, level):
    data = np.concatenate([l.reshape(-1) for l in levels])
    o = out3()
    L.or_cube_sample(data.ctypes.data, levels[0].shape[1], len(levels), d3(d), level, o)
    return np.array(list(o))


def test_cube_sample_texel_centres_midpoints_and_edges():
    rng = np.random.default_rng(5)
    size = 8
    cube = random_cube(rng, size)
    dirs = scene.cube_texel_dirs(size)
    for f in range(6):
        for y in range(size):
            for x in range(size):
                assert np.allclose(sample(cube, dirs[f, y, x], 0.0), cube[0][f, y, x, :3], atol=1e-6)
    # halfway between two texel centres of a row: their mean
    for f in range(6):
        y, x = 3, 4
        d = (dirs[f, y, x] / np.abs(dirs[f, y, x]).max() + dirs[f, y, x + 1] / np.abs(dirs[f, y, x + 1]).max()) / 2
        want = 0.5 * (cube[0][f, y, x, :3] + cube[0][f, y, x + 1, :3])
        assert np.allclose(sample(cube, d, 0.0), want, atol=1e-6)
    # beyond the last texel centre of a face: clamped to the edge texel
    face_pos_x = np.array([1.0, -0.999, -0.999])  # +x face, s and t near 1
    assert np.allclose(sample(cube, face_pos_x, 0.0), cube[0][0, size - 1, size - 1, :3], atol=1e-6)


def test_cube_levels_blend_linearly_and_clamp():
    mips = 5
    levels = [np.full((6, 16 >> m, 16 >> m, 4), float(m), np.float32) for m in range(mips)]
    for level in (0.0, 0.3, 1.0, 2.75, 4.0, 6.5, -1.0):
        got = sample(levels, (0.3, -0.5, 0.8), level)
        assert np.allclose(got, min(max(level, 0.0), mips - 1), atol=1e-12)


def test_cubemap_dir_is_the_quaternion_rotation_then_z_flip():
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(9)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        d = rng.normal(size=3)
        env = N.HmEnvProbe()
        env.view_rotation[:] = q.tolist()
        o = out3()
        L.or_cubemap_dir(C.byref(env), d3(d), o)
        want = Rotation.from_quat(q).apply(d)  # scipy quaternions are (x, y, z, w)
        want[2] = -want[2]
        assert np.allclose(list(o), want, atol=1e-12)


def probe(env, spec_levels, diffuse, tangent, view, albedo, roughness, tilt):
    spec = np.concatenate([l.reshape(-1) for l in spec_levels]).astype(np.float32)
    diff = np.ascontiguousarray(diffuse, np.float32)
    env.specular, env.specular_size, env.specular_mips = spec.ctypes.data, spec_levels[0].shape[1], len(spec_levels)
    env.diffuse, env.diffuse_size = diff.ctypes.data, diffuse.shape[1]
    o = out3()
    L.or_probe_radiance(C.byref(env), d3(tangent), d3(view), d3(albedo), roughness, tilt, o)
    return np.array(list(o))


def make_env(intensity=1.0, rotation=(0, 0, 0, 1)):
    env = N.HmEnvProbe()
    env.view_rotation[:] = list(rotation)
    env.intensity = intensity
    return env


def const_cube(size, rgb, mips=1):
    return [np.tile(np.array([*rgb, 1.0], np.float32), (6, max(size >> m, 1), max(size >> m, 1), 1))
            for m in range(mips)]


def test_probe_diffuse_of_grey_hair_is_albedo_times_irradiance():
    irr = (0.3, 0.5, 0.8)
    zero = const_cube(8, (0, 0, 0), 4)
    rng = np.random.default_rng(2)
    for _ in range(10):
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        a = rng.uniform(0.05, 0.9)
        got = probe(make_env(2.5), zero, const_cube(4, irr)[0], t, v, (a, a, a), 0.4, -0.05)
        # the luminance weights are f32 constants: their sum is 1 within 1e-7
        assert np.allclose(got, 2.5 * a * np.array(irr), rtol=1e-7)


def test_scatter_tint_special_cases():
    """scatter_tint (shade.wesl:81-85): no tint under full shadow, the square
    root of the hue (albedo over its Rec. 709 luminance) with none."""
    a = np.array([0.5, 0.2, 0.05])
    hue = a / (a @ np.array([0.2126, 0.7152, 0.0722]))
    o = out3()
    L.or_scatter_tint(d3(a), 1.0, o)
    assert np.allclose(list(o), 1.0, rtol=0, atol=0)
    L.or_scatter_tint(d3(a), 0.0, o)
    assert np.allclose(list(o), np.sqrt(hue), rtol=1e-6)


def test_probe_diffuse_tint_is_the_half_shadow_tint():
    """A coloured albedo under a grey irradiance: channel ratios are the
    albedo's times the tint's at shadow 0.5, (a_i / a_j)^(1 + 1/4)."""
    zero = const_cube(8, (0, 0, 0), 4)
    a = np.array([0.5, 0.2, 0.05])
    got = probe(make_env(), zero, const_cube(4, (1, 1, 1))[0], (1, 0, 0), (0, 0, 1), a, 0.4, 0.0)
    assert got[0] / got[1] == pytest.approx((a[0] / a[1]) ** 1.25, rel=1e-6)
    assert got[1] / got[2] == pytest.approx((a[1] / a[2]) ** 1.25, rel=1e-6)


def fresnel(c):
    return L.or_fresnel(c)


def test_probe_specular_weight_at_normal_incidence():
    f0 = ((1.55 - 1.0) / (1.55 + 1.0)) ** 2
    assert fresnel(1.0) == pytest.approx(f0, rel=1e-6)  # the ior is the f32 constant 1.55f
    f_trt = fresnel(0.5)
    # tangent perpendicular to the view: theta_o = 0, cos_o = 1
    t, v = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    zero_diff = const_cube(4, (0, 0, 0))[0]
    white = probe(make_env(), const_cube(8, (1, 1, 1), 4), zero_diff, t, v, (1, 1, 1), 0.3, 0.0)
    want = f0 + 0.76 * (1 - f_trt) ** 2 * f_trt  # white hair absorbs nothing
    assert np.allclose(white, want, rtol=1e-6)
    albedo, rough = (0.5, 0.25, 0.1), 0.35
    sig = (C.c_double * 3)()
    L.or_absorption_from_color(d3(albedo), rough, sig)
    got = probe(make_env(), const_cube(8, (1, 1, 1), 4), zero_diff, t, v, albedo, rough, 0.0)
    want = f0 + 0.76 * (1 - f_trt) ** 2 * f_trt * np.exp(-3.2 * np.array(list(sig)))
    assert np.allclose(got, want, rtol=1e-6)


def test_probe_specular_face_follows_the_tilt_and_rotation():
    # tangent +y, view +z: fiber normal +z, theta_o = 0, so the specular
    # direction is (0, sin tilt, cos tilt) before the z flip
    faces = [np.full((6, 8, 8, 4), 0.0, np.float32)]
    for f in range(6):
        faces[0][f, ..., :3] = f + 1.0
    zero_diff = const_cube(4, (0, 0, 0))[0]
    t, v = (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
    w = probe(make_env(), const_cube(8, (1, 1, 1)), zero_diff, t, v, (1, 1, 1), 0.3, 0.0)  # the weight
    assert probe(make_env(), faces, zero_diff, t, v, (1, 1, 1), 0.3, 0.0) == pytest.approx(6.0 * w)  # -z
    assert probe(make_env(), faces, zero_diff, t, v, (1, 1, 1), 0.3, -1.2) == pytest.approx(4.0 * w)  # -y
    assert probe(make_env(), faces, zero_diff, t, v, (1, 1, 1), 0.3, 1.2) == pytest.approx(3.0 * w)  # +y
    # a view off the fiber's normal plane: theta_o = 1, theta_i = tilt - 1
    # points down (face -y); theta_i = tilt + 1 would point up (face +y)
    v2 = (0.0, math.sin(1.0), math.cos(1.0))
    w2 = probe(make_env(), const_cube(8, (1, 1, 1)), zero_diff, t, v2, (1, 1, 1), 0.3, 0.0)
    assert probe(make_env(), faces, zero_diff, t, v2, (1, 1, 1), 0.3, 0.0) == pytest.approx(4.0 * w2)
    # a quarter turn about +y takes +z to +x, which the flip leaves +x
    q = (0.0, math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4))
    assert probe(make_env(rotation=q), faces, zero_diff, t, v, (1, 1, 1), 0.3, 0.0) == pytest.approx(1.0 * w)


def test_probe_roughness_selects_the_level():
    mips = 6
    levels = [np.full((6, 32 >> m, 32 >> m, 4), float(m), np.float32) for m in range(mips)]
    zero_diff = const_cube(4, (0, 0, 0))[0]
    t, v = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    w = probe(make_env(), const_cube(32, (1, 1, 1), mips), zero_diff, t, v, (1, 1, 1), 0.5, 0.0)
    for r in (0.04, 0.3, 0.55, 1.0, 1.3):
        got = probe(make_env(), levels, zero_diff, t, v, (1, 1, 1), r, 0.0)
        assert np.allclose(got, w * min(r * (mips - 1), mips - 1), rtol=1e-9)


# ---- rect lights through or_shade ------------------------------------------------

W, H = 32, 24


def one_pixel_frame():
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, W, H)
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 8, [])[0]
    x, y = 16, 12
    code = 0x820000
    word = L.or_pack_gb(code, 180, 90, 20, 30, 40, 50)
    center = np.full(W * H, EMPTY64, np.uint64)
    center[y * W + x] = word
    beta = np.full(W * H, 0xFFFFFFFF, np.uint32)
    beta[y * W + x] = (code << 8) | 200
    inp = (C.c_float * 16)()
    L.or_pixel_inputs(C.byref(cam), word, int(beta[y * W + x]), x, y, inp)
    return cam, center, beta, np.array(list(inp))[0:3], y * W + x


def shade(frame, points=(), rects=(), env=None):
    cam, center, beta, _, pix = frame
    pa, n_p, da, n_d = plugin.light_arrays(list(points), [])
    ra, n_r = plugin.rect_arrays(list(rects))
    dom = np.zeros(8 * 8 * 17, np.uint32)
    shaded = np.zeros((W * H, 4), np.float32)
    depth = np.zeros(W * H, np.float32)
    L.or_shade(C.byref(cam), center.ctypes.data, beta.ctypes.data, dom.ctypes.data, C.cast(pa, C.c_void_p), n_p,
               C.cast(da, C.c_void_p), n_d, C.cast(ra, C.c_void_p), n_r, env, 1.0, shaded.ctypes.data,
               depth.ctypes.data, 1, 1, 0, H)
    return shaded[pix, :3].astype(np.float64)


def lit_position(frame, dist=0.8):
    """The light position, among 200 directions around the pixel, that a
    point light lights most brightly (the hair lobes are narrow)."""
    world = frame[3]
    k = np.arange(200) + 0.5
    z = 1.0 - 2.0 * k / 200
    phi = np.pi * (1.0 + 5.0 ** 0.5) * k
    dirs = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z], 1)
    best, pos = -1.0, None
    for d in dirs:
        p = world + dist * d
        pt = {"position_radius": [*p.tolist(), 0.0], "color_inverse_square_range": [1.0, 1.0, 1.0, 0.04], "kind": 0}
        v = shade(frame, points=[pt]).sum()
        if v > best:
            best, pos = v, p
    assert best > 1e-2
    return pos


def test_small_rect_facing_the_pixel_is_a_point_light():
    frame = one_pixel_frame()
    world = frame[3]
    pos = lit_position(frame)
    rng_ = 5.0
    w, h, col = 0.01, 0.02, np.array([3.0, 2.0, 1.0])
    rect = scene.rect_facing(pos, world, w, h, col, rng_)
    point = {"position_radius": [*pos.tolist(), 0.0],
             "color_inverse_square_range": [*(col * w * h).tolist(), 1.0 / rng_ ** 2], "kind": 0}
    got, want = shade(frame, rects=[rect]), shade(frame, points=[point])
    assert want.min() > 1e-6
    assert np.allclose(got, want, rtol=1e-6, atol=0)  # the shaded buffer is f32


def test_rect_facing_away_or_out_of_range_adds_nothing_and_area_is_linear():
    frame = one_pixel_frame()
    world = frame[3]
    pos = lit_position(frame)
    away = scene.rect_facing(pos, 2 * pos - world, 0.2, 0.2, (3, 2, 1), 5.0)
    assert np.all(shade(frame, rects=[away]) == 0.0)
    far = scene.rect_facing(pos, world, 0.2, 0.2, (3, 2, 1), 0.5)  # range below the distance
    assert np.all(shade(frame, rects=[far]) == 0.0)
    one = shade(frame, rects=[scene.rect_facing(pos, world, 0.2, 0.1, (3, 2, 1), 5.0)])
    two = shade(frame, rects=[scene.rect_facing(pos, world, 0.2, 0.2, (3, 2, 1), 5.0)])
    assert one.min() > 1e-6 and np.allclose(two, 2.0 * one, rtol=1e-6, atol=0)


def test_probe_times_ao_in_the_pixel():
    """shade adds probe_radiance x ao: with the probe alone the pixel is the
    probe at the pixel's inputs times ao = 50/63 (the word's AO bits)."""
    frame = one_pixel_frame()
    cam, center, beta, _, pix = frame
    cam.flags[2] = 1
    env_l = scene.studio_environment(16, 4, 4)
    d = plugin.EnvProbeDesc(env_l)
    got = shade(frame, env=C.byref(d.desc))
    inp = (C.c_float * 16)()
    L.or_pixel_inputs(C.byref(cam), int(center[pix]), int(beta[pix]), pix % W, pix // W, inp)
    inp = np.array(list(inp), np.float64)
    o = out3()
    L.or_probe_radiance(C.byref(d.desc), d3(inp[3:6]), d3(inp[6:9]), d3(inp[9:12]), inp[12], float(cam.appearance[1]),
                        o)
    assert np.allclose(got, np.array(list(o)) * 50.0 / 63.0, rtol=1e-5)
    cam.flags[2] = 0  # AO off: ao = 1
    assert np.allclose(shade(frame, env=C.byref(d.desc)), np.array(list(o)), rtol=1e-5)
