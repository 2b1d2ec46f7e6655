"""Small scenes that run the branches of hair_pass the measured workloads do
not: spot lights (both signs of the stored direction's y), several directional
lights (the first shadow caster takes the DOM, shade.wesl:482-498), no shadow
caster at all, a forward-Z camera with scene depth (pass.rs:778-787,
raster.wgsl:93-99), orthographic views (shade.wesl:560-570), and the HairGroom
toggles lod / ambient_occlusion / deep_opacity off (lib.rs:45-64).

Each case is the c1 groom (4096 straight strands x 16 segments) at 480x270,
DOM 512^2, rendered the same way by the HIP path and the oracle.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from bevy_hair_amd import plugin, scene
from bevy_hair_amd.groom import synthetic_groom

WIDTH, HEIGHT, DOM = 480, 270, 512


def _dir(v, color, casts):
    d = np.asarray(v, dtype=np.float64)
    d /= np.linalg.norm(d)
    return {"direction_to_light": [*d.tolist(), 0.0], "color": [*color, 1.0], "flags": 1 if casts else 0}


@dataclasses.dataclass
class Case:
    name: str
    mesh: object
    view: scene.View
    groom: plugin.HairGroom
    points: list
    dirs: list
    params: object
    control_points: int
    strand_fraction: float
    scene_depth: np.ndarray | None
    scene_clear: float

    def light_arrays(self):
        return plugin.light_arrays(self.points, self.dirs)


CASE_NAMES = ["spot", "directionals", "no_caster", "forward_z", "orthographic", "orthographic_forward_z",
              "lod_off", "ao_off", "dom_off"]


def _checker(view: scene.View, mesh, clear: float, cell: int = 12) -> np.ndarray:
    """Scene depth: every other `cell`-pixel square at the clip depth of the
    groom's bounds centre (occluding part of the groom), the rest `clear`."""
    c = np.append(0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64), 1.0)
    clip = view.clip_from_world @ c
    z = np.float32(clip[2] / max(clip[3], 1e-5))
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    on = ((yy // cell + xx // cell) % 2) == 0
    return np.where(on, z, np.float32(clear)).astype(np.float32).reshape(-1)


def _with_ao(mesh):
    """The groom with a root-to-tip ambient occlusion ramp in its cage
    (pos_ao.w of layer l: 0.35 + 0.65 sqrt(l / 15)); the synthetic generator
    writes AO = 1, which leaves the G-buffer's AO field constant."""
    layer = (np.arange(len(mesh.layers)) % 64) // 4
    mesh.layers[:, 3] = (0.35 + 0.65 * np.sqrt(layer / 15.0)).astype(np.float32)
    return mesh


def make_case(name: str) -> Case:
    mesh = _with_ao(synthetic_groom("straight", 4096, 1))
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max).astype(np.float64)
    rng = np.random.default_rng(77)
    points = scene.point_lights(rng, 2, centre)
    dirs = [scene.key_light()]
    projection = "perspective"
    groom_kw = {}
    control_points, fraction = 17, 1.0
    scene_depth, clear = None, 0.0
    if name == "spot":
        # one spot from above (y < 0: the flag), one from below (y > 0), aimed
        # off-centre so the cone edges cross the groom, bright against a dim
        # key and point light: they light ~40 % of the hit pixels, ~4 % in
        # the cone falloff
        points = points[:1]
        points += [scene.spot_light(centre + [0.3, 1.2, 0.6], centre + [0.05, 0.0, 0.0], 20.0, 40.0, (15.0, 12.0, 10.0)),
                   scene.spot_light(centre + [0.8, -0.9, 0.5], centre + [0.0, 0.1, 0.0], 14.0, 28.0, (6.0, 9.0, 14.0))]
        dirs = [scene.key_light(color=(0.3, 0.3, 0.3))]
    elif name == "directionals":
        # first light casts no shadow, the second is the key (DOM) light, the
        # third also casts but gets no DOM depth
        dirs = [_dir((-0.3, 0.8, 0.5), (0.8, 0.8, 1.0), False), scene.key_light(),
                _dir((0.2, -0.4, -0.9), (0.7, 0.5, 0.4), True)]
    elif name == "no_caster":
        dirs = [_dir((-0.3, 0.8, 0.5), (0.8, 0.8, 1.0), False), _dir((0.6, 0.3, 0.7), (1.2, 1.1, 1.0), False)]
    elif name == "forward_z":
        projection = "perspective_forward_z"
        clear = 1.0
    elif name == "orthographic":
        projection = "orthographic"
    elif name == "orthographic_forward_z":
        projection = "orthographic_forward_z"
        clear = 1.0
    elif name == "lod_off":
        groom_kw["lod"] = False
        control_points, fraction = 0, 0.0  # the reference's LOD record: (N_b, 127) with LOD off
    elif name == "ao_off":
        groom_kw["ambient_occlusion"] = False
    elif name == "dom_off":
        groom_kw["deep_opacity"] = False
    else:
        raise ValueError(name)
    view = scene.fit_camera(mesh, WIDTH, HEIGHT, projection=projection)
    if name in ("forward_z", "orthographic_forward_z"):
        scene_depth = _checker(view, mesh, clear)
    groom = plugin.HairGroom(**groom_kw)
    params = plugin.Renderer.make_params(view, groom, mesh, DOM, dirs)
    return Case(name, mesh, view, groom, points, dirs, params, control_points, fraction, scene_depth, clear)


def oracle_frame(case: Case, threads: int = 1):
    """The oracle's hair_pass for the case (oracle imported here: test use only)."""
    import oracle

    ref = oracle.Frame(case.mesh, WIDTH, HEIGHT, DOM)
    ref.scene_depth = case.scene_depth
    ref.scene_clear = case.scene_clear
    pa, n_p, da, n_d = case.light_arrays()
    ref.hair_pass(case.params, pa, n_p, da, n_d, control_points=case.control_points,
                  strand_fraction=case.strand_fraction, threads=threads)
    return ref
