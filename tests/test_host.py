"""Host logic and the C ABI without a GPU: symbol export, parameter
construction (pass.rs:399-453, 676-787), key-light choice, HAIRBAK1 cache."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from bevy_hair_amd import _native as N
from bevy_hair_amd import groom as G
from bevy_hair_amd import plugin, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions() -> set[str]:
    text = open(os.path.join(ROOT, "include", "hair_mi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol(native_lib):
    declared = header_functions()
    assert declared == set(N.ABI), f"binding and header disagree: {declared ^ set(N.ABI)}"
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (hm_[a-z0-9_]+)", out.stdout))
    assert declared <= exported, f"missing exports: {declared - exported}"


def test_struct_sizes_match_reference_layouts():
    assert C.sizeof(N.HmParams) == 432  # HAIR_PARAMS_SIZE, pass.rs:62
    assert C.sizeof(N.HmShadowLight) == 112  # SHADOW_LIGHT_BYTES, pass.rs:29
    offs = {f: getattr(N.HmParams, f).offset for f, _ in N.HmParams._fields_}
    assert offs == {"clip_from_world": 0, "world_from_clip": 64, "world_from_model": 128,
                    "light_clip_from_world": 192, "camera_pos": 256, "camera_forward": 272, "screen": 288,
                    "appearance": 304, "filter_params": 320, "albedo": 336, "flags": 352, "pass_mode": 368,
                    "light_eye": 384, "light_forward": 400, "dom_info": 416}


def test_create_without_gpu_fails_cleanly(native_lib):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    h = C.c_void_p()
    assert native_lib.hm_create(0, C.byref(h)) == 4  # HM_ERR_NO_DEVICE, no abort
    assert not h.value


def test_status_strings(native_lib):
    assert native_lib.hm_status_string(0) == b"ok"
    assert native_lib.hm_status_string(1) == b"invalid argument"


def m4(arr):
    return np.array(list(arr), dtype=np.float64).reshape(4, 4).T  # column-major -> rows


@pytest.fixture(scope="module")
def c1():
    from tests.scenes import make_scene
    return make_scene("c1")


def test_make_params_camera_record(c1):
    cam, dom1, dom2 = c1.params[0], c1.params[1], c1.params[2]
    w = c1.w
    assert list(cam.screen)[:3] == [w.width, w.height, np.float32(0.05)]
    lo, hi = c1.mesh.bounds_min.astype(np.float64), c1.mesh.bounds_max.astype(np.float64)
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    centre = corners.mean(0)
    radius = max(1.0, np.linalg.norm(corners - centre, axis=1).max())
    eye = c1.view.world_from_view[:3, 3]
    dist = max(np.linalg.norm(eye - centre), 0.5)
    assert abs(cam.screen[3] - max(dist + 4 * radius, 1.05)) < 1e-4  # cam_far, pass.rs:409
    assert list(cam.pass_mode) == [0, c1.mesh.bundle_count, 1, 0]  # reverse-Z detected
    assert list(dom1.pass_mode) == [1, c1.mesh.bundle_count, 1, 1]
    assert list(dom2.pass_mode) == [2, c1.mesh.bundle_count, 1, 1]
    assert list(cam.flags) == [1, 1, 1, 1]
    assert np.allclose(list(cam.appearance), [0.2, 0.08, 3.0, 1.0])
    fp = list(cam.filter_params)
    assert np.allclose(fp[:2], [1.0, 0.2]) and np.isclose(fp[3], 0.05)
    assert np.isclose(fp[2], (dom1.dom_info[3] - dom1.dom_info[2]) / 16.0)  # dom_scale, pass.rs:411
    assert list(cam.dom_info)[:2] == [w.dom_size, w.dom_size]
    # camera basis from world_from_view (pass.rs:342-344)
    assert np.allclose(list(cam.camera_pos)[:3], eye, atol=1e-5) and cam.camera_pos[3] == 1.0
    fwd = -c1.view.world_from_view[:3, 2]
    assert np.allclose(list(cam.camera_forward)[:3], fwd, atol=1e-6) and cam.camera_forward[3] == 0.0
    # world_from_clip inverts clip_from_world
    assert np.allclose(m4(cam.world_from_clip) @ m4(cam.clip_from_world), np.eye(4), atol=1e-4)
    # albedo: sRGB (0.16, 0.07, 0.035) to linear (lib.rs:58, mod.rs:126)
    assert np.allclose(list(cam.albedo), [plugin.srgb_to_linear(c) for c in (0.16, 0.07, 0.035)] + [1.0])


def test_light_fit_frames_the_groom(c1):
    cam, dom1 = c1.params[0], c1.params[1]
    lo, hi = c1.mesh.bounds_min.astype(np.float64), c1.mesh.bounds_max.astype(np.float64)
    corners = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    clip = corners @ m4(dom1.clip_from_world).T
    ndc = clip[:, :2] / clip[:, 3:]
    # the orthographic light view spans [-radius, radius] around the bounds centre,
    # so every corner of the AABB lands inside the map (pass.rs:747-776)
    assert (np.abs(ndc) <= 1.0 + 1e-5).all()
    key = np.array(scene.SUN_DIR)
    assert np.allclose(list(dom1.camera_forward)[:3], -key, atol=1e-6)
    assert np.allclose(list(cam.light_forward)[:3], -key, atol=1e-6)
    # the eye sits one radius toward the light, so no corner is behind it; every
    # corner is before the far plane, and near only clamps up to 0.05 (pass.rs:763-764)
    eye = np.array(list(dom1.light_eye)[:3])
    depth = (corners[:, :3] - eye) @ -key
    near, far = dom1.dom_info[2], dom1.dom_info[3]
    assert (depth >= -1e-5).all() and (depth < far).all()
    assert near <= max(depth.min(), 0.05) + 1e-6 and far >= near + 1.0 - 1e-6
    assert list(dom1.screen) == [c1.w.dom_size, c1.w.dom_size, near, far]
    assert np.allclose(m4(dom1.light_clip_from_world), m4(cam.light_clip_from_world))


def test_key_direction_rules(native_lib):
    def key(lights):
        arr = (N.HmDirectionalLight * max(1, len(lights)))(*[plugin.directional_light(l) for l in lights])
        out = (C.c_float * 3)()
        native_lib.hm_key_direction(arr, len(lights), out)
        return np.array(list(out))

    d = np.array([0.35, 0.82, 0.45])
    assert np.allclose(key([]), d / np.linalg.norm(d))  # pass.rs:717
    a = {"direction_to_light": [0, 0, 2, 0], "color": [1, 1, 1, 1], "flags": 0}
    b = {"direction_to_light": [3, 0, 0, 0], "color": [1, 1, 1, 1], "flags": 1}
    assert np.allclose(key([a]), [0, 0, 1])  # first light when none casts
    assert np.allclose(key([a, b]), [1, 0, 0])  # first shadow caster wins


def test_shadow_light_record(native_lib, c1):
    out = N.HmShadowLight()
    native_lib.hm_shadow_light_from_params(C.byref(c1.params[1]), C.byref(out))
    assert list(out.light_clip_from_world) == list(c1.params[1].clip_from_world)
    assert list(out.depth) == [c1.params[1].dom_info[2], c1.params[1].dom_info[3], np.float32(0.05), 1.0]
    assert out.light_eye[3] == 1.0 and out.light_forward[3] == 0.0


# ---- groom cache --------------------------------------------------------------------------


def test_hairbak1_roundtrip_and_rejections(tmp_path):
    mesh = G.synthetic_groom("straight", 600, 5)
    stamp = G.SourceStamp(1234, 1700000000, 5, 99, 1600000000, 7)
    blob = G.encode_cache(stamp, mesh)
    assert blob[:8] == b"HAIRBAK1"
    assert len(blob) == 92 + mesh.bundle_count * 12 + len(mesh.layers) * 28 + len(mesh.style) * 16
    back = G.decode_cache(blob, stamp)
    assert back is not None
    assert np.array_equal(back.bundles, mesh.bundles) and np.array_equal(back.layers[:, :7], mesh.layers[:, :7])
    assert np.array_equal(back.style, mesh.style)
    assert np.array_equal(back.bounds_min, mesh.bounds_min) and np.array_equal(back.bounds_max, mesh.bounds_max)
    # examples/bake.rs:188-212: magic, fingerprint, stamps and payload length all gate the cache
    assert G.decode_cache(b"HAIRBAK2" + blob[8:], stamp) is None
    assert G.decode_cache(blob, stamp, fingerprint=G.BAKE_FINGERPRINT ^ 1) is None
    assert G.decode_cache(blob, G.SourceStamp(1235, 1700000000, 5, 99, 1600000000, 7)) is None
    assert G.decode_cache(blob[:-4], stamp) is None
    assert G.decode_cache(blob + b"\0", stamp) is None
    path = tmp_path / "cache" / "g.bake"
    G.write_cache(str(path), stamp, mesh)
    assert not os.path.exists(str(path) + ".tmp")
    assert G.read_cache(str(path), stamp) is not None
    assert G.read_cache(str(tmp_path / "missing.bake"), stamp) is None


def test_synthetic_grooms_are_valid_and_seeded():
    a = G.synthetic_groom("wavy", 3000, 9)
    a.validate()
    assert a.strand_count == 3000
    assert np.isfinite(a.layers).all() and np.isfinite(a.style).all()
    b = G.bake(G.synth_strands("wavy", 3000, 9))
    assert np.array_equal(a.layers, b.layers) and np.array_equal(a.bundles, b.bundles)
    for kind in ("straight", "curly"):
        G.synthetic_groom(kind, 500, 1).validate()
