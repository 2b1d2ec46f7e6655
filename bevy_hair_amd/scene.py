"""Views, lights and the seeded workloads the path is measured on.

Bevy supplies the camera and light list in the reference (ExtractedView,
MeshViewBindGroup); here they are built explicitly:
  - camera: Transform::looking_at + perspective_infinite_reverse_rh (Bevy's
    default PerspectiveProjection, near 0.1) with the example's 50 mm lens,
    fov = 2 atan(12 / 50) (examples/groom.rs:252-255);
  - key light: the example's sun direction (0.788, 0.217, 0.576)
    (examples/groom.rs:240), shadow-casting so the DOM path is active.
Light colours are linear radiance scales with exposure 1 (Bevy's photometric
units and exposure are outside the hair path).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import groom as groom_mod

FOV_Y = 2.0 * np.arctan(24.0 / (2.0 * 50.0))
SUN_DIR = np.array([0.788, 0.217, 0.576]) / np.linalg.norm([0.788, 0.217, 0.576])


def look_at_world_from_view(eye, focus, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Transform::from_translation(eye).looking_at(focus, up) as a column-major matrix."""
    eye, focus, up = (np.asarray(v, dtype=np.float64) for v in (eye, focus, up))
    back = eye - focus
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    upv = np.cross(back, right)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, upv, back, eye
    return m


def view_from_world_rigid(world_from_view: np.ndarray) -> np.ndarray:
    """Inverse of a rigid transform as R^T, -R^T t, with the last row exactly
    (0, 0, 0, 1) (a general inverse leaves rounding residue there, which the
    orthographic test of shade.wesl:560 would read as a perspective row)."""
    r, t = world_from_view[:3, :3], world_from_view[:3, 3]
    m = np.eye(4)
    m[:3, :3] = r.T
    m[:3, 3] = -r.T @ t
    return m


def perspective_rh(fov_y: float, aspect: float, near: float, far: float) -> np.ndarray:
    """glam Mat4::perspective_rh: forward Z (near -> 0, far -> 1), finite far."""
    h = 1.0 / np.tan(0.5 * fov_y)
    r = far / (near - far)
    m = np.zeros((4, 4))
    m[0, 0] = h / aspect
    m[1, 1] = h
    m[2, 2] = r
    m[2, 3] = r * near
    m[3, 2] = -1.0
    return m


def orthographic_rh(left: float, right: float, bottom: float, top: float, near: float, far: float) -> np.ndarray:
    """glam Mat4::orthographic_rh (z: near -> 0, far -> 1). Bevy's
    OrthographicProjection passes (far, near) for a reverse-Z depth."""
    rw, rh, r = 1.0 / (right - left), 1.0 / (top - bottom), 1.0 / (near - far)
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = 2.0 * rw, 2.0 * rh, r
    m[0, 3], m[1, 3], m[2, 3] = -(left + right) * rw, -(top + bottom) * rh, r * near
    return m


def perspective_infinite_reverse_rh(fov_y: float, aspect: float, near: float) -> np.ndarray:
    f = 1.0 / np.tan(0.5 * fov_y)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[3, 2] = -1.0
    m[2, 3] = near
    return m


def cols(m: np.ndarray) -> list[float]:
    """Row-major numpy 4x4 -> column-major float list (glam to_cols_array)."""
    return [float(x) for x in np.asarray(m, dtype=np.float32).T.reshape(-1)]


@dataclasses.dataclass
class View:
    clip_from_world: np.ndarray  # 4x4, numpy row/col convention (m[row, col])
    world_from_view: np.ndarray
    width: int
    height: int


PROJECTIONS = ("perspective", "perspective_forward_z", "orthographic", "orthographic_forward_z")


def fit_camera(mesh: groom_mod.HairMesh, width: int, height: int, heading=(150.0, 0.0, 40.0),
               pitch_deg: float = 15.0, margin: float = 1.1, projection: str = "perspective") -> View:
    """Looks at the groom's bounds centre from a distance that fits its bounding sphere.

    projection: "perspective" (Bevy's default: infinite reverse-Z, near 0.1),
    "perspective_forward_z" (finite forward-Z, near 0.1, far 100),
    "orthographic" (Bevy's OrthographicProjection: reverse-Z, near 0.1,
    far 100, the sphere's diameter tall) or "orthographic_forward_z"."""
    lo, hi = mesh.bounds_min.astype(np.float64), mesh.bounds_max.astype(np.float64)
    focus = 0.5 * (lo + hi)
    radius = 0.5 * np.linalg.norm(hi - lo)
    h = np.asarray(heading, dtype=np.float64)
    h /= np.linalg.norm(h)
    pitch = np.radians(pitch_deg)
    dist = radius * margin / np.sin(0.5 * FOV_Y)
    eye = focus + h * (dist * np.cos(pitch)) + np.array([0.0, 1.0, 0.0]) * (dist * np.sin(pitch))
    wfv = look_at_world_from_view(eye, focus)
    aspect = width / height
    if projection == "perspective":
        proj = perspective_infinite_reverse_rh(FOV_Y, aspect, 0.1)
        return View(proj @ np.linalg.inv(wfv), wfv, width, height)
    half = radius * margin
    if projection == "perspective_forward_z":
        proj = perspective_rh(FOV_Y, aspect, 0.1, 100.0)
    elif projection == "orthographic":
        proj = orthographic_rh(-half * aspect, half * aspect, -half, half, 100.0, 0.1)
    elif projection == "orthographic_forward_z":
        proj = orthographic_rh(-half * aspect, half * aspect, -half, half, 0.1, 100.0)
    else:
        raise ValueError(f"unknown projection {projection!r}")
    return View(proj @ view_from_world_rigid(wfv), wfv, width, height)


def spot_light(position, target, inner_deg: float, outer_deg: float, color, light_range: float = 3.0) -> dict:
    """A spot light as Bevy's clustered light record holds it (shade.wesl:312-327):
    light_custom_data = (dir.x, dir.z, spot scale, spot offset) of the unit
    direction the light points along, the sign of dir.y in the flags, scale
    1 / max(cos inner - cos outer, 1e-4) and offset -cos outer * scale."""
    position = np.asarray(position, dtype=np.float64)
    d = np.asarray(target, dtype=np.float64) - position
    d /= np.linalg.norm(d)
    ci, co = np.cos(np.radians(inner_deg)), np.cos(np.radians(outer_deg))
    scale = 1.0 / max(ci - co, 1e-4)
    flags = 2 if d[1] < 0.0 else 0  # HM_LIGHT_FLAG_SPOT_Y_NEGATIVE
    return {"position_radius": [*position.tolist(), 0.0],
            "color_inverse_square_range": [*[float(c) for c in color], 1.0 / (light_range * light_range)],
            "light_custom_data": [float(d[0]), float(d[2]), float(scale), float(-co * scale)],
            "flags": flags, "kind": 1}


@dataclasses.dataclass
class Lights:
    points: np.ndarray  # structured by _native.HmPointLight fields, as a list of dicts
    directionals: list


def key_light(color=(3.0, 2.9, 2.7)) -> dict:
    return {"direction_to_light": [*SUN_DIR.tolist(), 0.0], "color": [*color, 1.0], "flags": 1}


def point_lights(rng: np.random.Generator, count: int, centre: np.ndarray, box: float = 2.0,
                 light_range: float = 3.0, intensity: float = 0.6) -> list[dict]:
    """`count` point lights uniform in a box of side `box` around `centre`."""
    out = []
    for _ in range(count):
        pos = centre + rng.uniform(-0.5 * box, 0.5 * box, 3)
        col = rng.uniform(0.3, 1.0, 3) * intensity
        out.append({"position_radius": [*pos.tolist(), 0.0],
                    "color_inverse_square_range": [*col.tolist(), 1.0 / (light_range * light_range)],
                    "light_custom_data": [0.0, 0.0, 0.0, 0.0], "flags": 0, "kind": 0})
    return out


@dataclasses.dataclass
class Workload:
    """One measured configuration: groom, target sizes, lights and LOD override."""

    name: str
    kind: str
    strands: int
    seed: int
    width: int
    height: int
    dom_size: int
    control_points: int  # camera ncp override (0 = reference LOD)
    strand_fraction: float
    n_point_lights: int
    albedo_linear: tuple | None = None
    filter: bool = True


WORKLOADS = {
    # 4096 straight strands x 16 segments, 640x360, one directional, DOM 512^2.
    "c1": Workload("c1_straight_4096x16_640x360", "straight", 4096, 1, 640, 360, 512, 17, 1.0, 0),
    # 100k wavy strands x 64 segments, 1080p, DOM 2048^2, key + 32 point lights.
    "c2": Workload("c2_wavy_100000x64_1920x1080", "wavy", 100_000, 2, 1920, 1080, 2048, 65, 1.0, 32),
    # 60k curly strands x 126 segments (ncp 127 is the maximum), 2160p, key + 128 lights.
    "c3": Workload("c3_curly_60000x126_3840x2160", "curly", 60_000, 3, 3840, 2160, 2048, 127, 1.0, 128),
    # The learned shading term's evaluation: the c1 groom rendered at 1080p.
    "c4": Workload("c4_straight_4096x16_1920x1080", "straight", 4096, 1, 1920, 1080, 512, 17, 1.0, 0),
    # Small parity cases.
    "tiny": Workload("tiny_straight_512x16_128x96", "straight", 512, 7, 128, 96, 128, 17, 0.0, 2),
    "small_lod": Workload("small_wavy_2048_lod_256x160", "wavy", 2048, 11, 256, 160, 256, 0, 0.0, 4),
}


def workload_lights(w: Workload, mesh: groom_mod.HairMesh):
    rng = np.random.default_rng(1000 + w.seed)
    centre = 0.5 * (mesh.bounds_min.astype(np.float64) + mesh.bounds_max.astype(np.float64))
    pts = point_lights(rng, w.n_point_lights, centre)
    albedo = w.albedo_linear
    if albedo is None and w.n_point_lights >= 32:
        albedo = tuple(rng.uniform(0.05, 0.6, 3).tolist())
    return pts, [key_light()], albedo
