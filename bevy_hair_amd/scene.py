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


def rect_light(position, right, up, width: float, height: float, color, light_range: float = 3.0) -> dict:
    """A rectangular area light (shade.wesl:503-518): centre, the unit right and
    up directions of the rectangle (it emits along normalize(up x right)),
    its size, linear colour and range."""
    return {"position": [*[float(c) for c in position], 0.0], "right": [*[float(c) for c in right], 0.0],
            "up": [*[float(c) for c in up], 0.0], "color": [*[float(c) for c in color], 1.0],
            "width": float(width), "height": float(height), "range": float(light_range)}


def rect_facing(position, target, width: float, height: float, color, light_range: float = 3.0) -> dict:
    """A rect light at `position` whose emitting side faces `target`."""
    position = np.asarray(position, np.float64)
    n = np.asarray(target, np.float64) - position
    n /= np.linalg.norm(n)
    helper = np.array([0.0, 1.0, 0.0]) if abs(n[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    right = np.cross(helper, n)
    right /= np.linalg.norm(right)
    up = np.cross(n, right)  # up x right = n
    return rect_light(position, right, up, width, height, color, light_range)


@dataclasses.dataclass
class EnvironmentMapLight:
    """The view's environment probe (Bevy's EnvironmentMapLight as shade.wesl
    reads it): a prefiltered specular cube map with its mip chain, a diffuse
    (irradiance / pi) cube map, a rotation quaternion (x, y, z, w) and an
    intensity. Each cube map level is float32 [6, size, size, 4] (RGBA), faces
    +x, -x, +y, -y, +z, -z, rows from the top of each face."""

    specular: list  # level 0 first; level m has size max(S >> m, 1)
    diffuse: np.ndarray
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    intensity: float = 1.0

    def packed(self):
        spec = np.concatenate([np.ascontiguousarray(l, np.float32).reshape(-1) for l in self.specular])
        diff = np.ascontiguousarray(self.diffuse, np.float32).reshape(-1)
        return spec, diff

    def validate(self):
        S = self.specular[0].shape[1]
        for m, l in enumerate(self.specular):
            sm = max(S >> m, 1)
            if l.shape != (6, sm, sm, 4):
                raise ValueError(f"specular level {m}: shape {l.shape}, expected {(6, sm, sm, 4)}")
        D = self.diffuse.shape[1]
        if self.diffuse.shape != (6, D, D, 4):
            raise ValueError(f"diffuse shape {self.diffuse.shape}")


def cube_texel_dirs(size: int, edge_snap: bool = False) -> np.ndarray:
    """Unit directions [6, size, size, 3] through the texel centres of a cube
    map (Vulkan's face orientation). With edge_snap the border texels take
    the direction of the point on the face edge (and corner) beside them, so a
    map filled from a continuous function matches across face seams."""
    c = (np.arange(size) + 0.5) / size
    if edge_snap and size > 1:
        c[0], c[-1] = 0.0, 1.0
    sc = 2.0 * c[None, :] - 1.0 + 0.0 * c[:, None]   # [t, s]
    tc = 2.0 * c[:, None] - 1.0 + 0.0 * c[None, :]
    one = np.ones_like(sc)
    faces = [np.stack([one, -tc, -sc], -1), np.stack([-one, -tc, sc], -1),
             np.stack([sc, one, tc], -1), np.stack([sc, -one, -tc], -1),
             np.stack([sc, -tc, one], -1), np.stack([-sc, -tc, -one], -1)]
    d = np.stack(faces)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def studio_environment(size: int = 64, mips: int = 6, diffuse_size: int = 16, intensity: float = 1.0,
                       rotation=(0.0, 0.0, 0.0, 1.0)) -> EnvironmentMapLight:
    """A procedural studio probe: a sky-to-floor gradient with two soft
    boxes, each level filled from a smooth function of direction whose lobes
    widen with the level (a stand-in for a GGX prefilter), the diffuse map
    from a wide version. Seam-continuous (cube_texel_dirs edge_snap)."""
    def radiance(d, width):
        sky = np.array([0.55, 0.62, 0.75]) * (0.5 + 0.5 * d[..., 1:2]) + np.array([0.12, 0.1, 0.08]) * (
            0.5 - 0.5 * d[..., 1:2])
        out = sky
        for axis, col in ((np.array([0.6, 0.7, 0.38]), np.array([6.0, 5.6, 5.0])),
                          (np.array([-0.8, 0.3, -0.52]), np.array([1.5, 2.0, 3.0]))):
            a = axis / np.linalg.norm(axis)
            cosang = np.clip(d @ a, -1.0, 1.0)[..., None]
            out = out + col * np.exp((cosang - 1.0) / (0.02 + width))
        return out

    levels = []
    for m in range(mips):
        sm = max(size >> m, 1)
        d = cube_texel_dirs(sm, edge_snap=True)
        rgb = radiance(d, 0.08 * m * m)
        levels.append(np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,))], -1).astype(np.float32))
    d = cube_texel_dirs(diffuse_size, edge_snap=True)
    rgb = radiance(d, 2.0) * 0.35
    diffuse = np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,))], -1).astype(np.float32)
    return EnvironmentMapLight(levels, diffuse, tuple(rotation), intensity)


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


def scattered_lights(rng: np.random.Generator, count: int, centre, radius: float, range_lo: float,
                     range_hi: float, intensity: float = 0.05) -> list[dict]:
    """`count` short-range lights, point and spot alternating, uniform in a
    box 3.2 radii wide around `centre`, ranges U(range_lo, range_hi) x radius:
    each reaches a small part of a groom of that radius."""
    centre = np.asarray(centre, np.float64)
    out = []
    for i in range(count):
        pos = centre + rng.uniform(-1.6, 1.6, 3) * radius
        col = rng.uniform(0.2, 1.0, 3) * intensity
        rng_ = float(rng.uniform(range_lo, range_hi) * radius)
        if i % 2:
            out.append(spot_light(pos, centre + rng.uniform(-0.3, 0.3, 3) * radius, 15.0, 40.0, col, rng_))
        else:
            out.append({"position_radius": [*pos.tolist(), 0.0],
                        "color_inverse_square_range": [*col.tolist(), 1.0 / (rng_ * rng_)],
                        "light_custom_data": [0.0, 0.0, 0.0, 0.0], "flags": 0, "kind": 0})
    return out


def studio_rects(centre, radius: float) -> list[dict]:
    """Three rect lights around a groom: a warm key panel, a cool fill and a
    short-range red kicker below."""
    centre = np.asarray(centre, np.float64)
    return [rect_facing(centre + np.array([1.2, 0.8, 0.9]) * radius, centre, 0.3 * radius, 0.2 * radius,
                        (4.0, 3.5, 3.0), 6.0 * radius),
            rect_facing(centre + np.array([-1.5, 0.2, 0.6]) * radius, centre, 0.5 * radius, 0.5 * radius,
                        (0.5, 1.0, 2.0), 6.0 * radius),
            rect_facing(centre + np.array([0.0, -1.4, -0.9]) * radius, centre, 0.2 * radius, 0.4 * radius,
                        (2.0, 0.5, 0.5), 1.0 * radius)]


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
    # 20k wavy strands on a 64^2 DOM: tens of thousands of samples per 16 x 16
    # tile, more than a DOM tile workgroup keeps between its two passes.
    "dense_dom": Workload("dense_dom_wavy_20000x64_320x180", "wavy", 20_000, 5, 320, 180, 64, 65, 1.0, 1),
}


def workload_lights(w: Workload, mesh: groom_mod.HairMesh):
    rng = np.random.default_rng(1000 + w.seed)
    centre = 0.5 * (mesh.bounds_min.astype(np.float64) + mesh.bounds_max.astype(np.float64))
    pts = point_lights(rng, w.n_point_lights, centre)
    albedo = w.albedo_linear
    if albedo is None and w.n_point_lights >= 32:
        albedo = tuple(rng.uniform(0.05, 0.6, 3).tolist())
    return pts, [key_light()], albedo
