"""A ready-to-render workload: seeded groom, fitted camera, lights and HairParams."""
from __future__ import annotations

import dataclasses

import numpy as np

from . import plugin, scene
from .groom import HairMesh, synthetic_groom


@dataclasses.dataclass
class Scene:
    w: scene.Workload
    mesh: HairMesh
    view: scene.View
    groom: plugin.HairGroom
    points: list
    dirs: list
    params: object  # (HmParams * 3): camera, DOM depth, DOM opacity

    def light_arrays(self):
        return plugin.light_arrays(self.points, self.dirs)


def make_scene(name: str, **groom_kw) -> Scene:
    """The named workload of scene.WORKLOADS with HairGroom defaults (lib.rs:54-64) unless overridden."""
    w = scene.WORKLOADS[name]
    mesh = synthetic_groom(w.kind, w.strands, w.seed)
    view = scene.fit_camera(mesh, w.width, w.height)
    points, dirs, albedo = scene.workload_lights(w, mesh)
    kw = dict(groom_kw)
    if albedo is not None and "albedo_linear" not in kw:
        kw["albedo_linear"] = albedo
    kw.setdefault("filter", w.filter)
    groom = plugin.HairGroom(**kw)
    params = plugin.Renderer.make_params(view, groom, mesh, w.dom_size, dirs)
    return Scene(w, mesh, view, groom, points, dirs, params)


# The batched multi-view frame: 24 views over three grooms of uneven size at
# mixed resolutions. (kind, strands, seed, camera ncp): 20k x 16, 100k x 64 and
# 150k x 126 segments (126 is the most a strand can have: ncp 127).
BATCH_GROOMS = [("straight", 20_000, 21, 17), ("wavy", 100_000, 2, 65), ("curly", 150_000, 23, 127)]
BATCH_SIZES = [(1280, 720), (1920, 1080), (2560, 1440)]
BATCH_VIEWS = 24
BATCH_DOM = 2048


@dataclasses.dataclass
class BatchScene:
    meshes: list          # one HairMesh per groom
    grooms: list          # one HairGroom per groom
    views: list           # (groom index, width, height, View, params, camera ncp)
    points: list
    dirs: list


def make_batch_scene(views: int = BATCH_VIEWS, dom_size: int = BATCH_DOM) -> BatchScene:
    """`views` views, groom g = v % 3, size v % 3 rotated by v // 3, camera
    heading turned 360 / (views / 3) degrees per view of a groom; the c2 light
    set (key + 32 point lights) around the 100k groom."""
    meshes = [synthetic_groom(kind, n, seed) for kind, n, seed, _ in BATCH_GROOMS]
    grooms = [plugin.HairGroom(filter=True) for _ in meshes]
    points, dirs, albedo = scene.workload_lights(scene.WORKLOADS["c2"], meshes[1])
    grooms[1] = plugin.HairGroom(filter=True, albedo_linear=albedo)
    per = max(1, views // len(meshes))
    out = []
    for v in range(views):
        g = v % len(meshes)
        w, h = BATCH_SIZES[(v + v // len(meshes)) % len(BATCH_SIZES)]
        ang = 2.0 * np.pi * (v // len(meshes)) / per
        heading = (np.cos(ang) * 150.0 - np.sin(ang) * 40.0, 0.0, np.sin(ang) * 150.0 + np.cos(ang) * 40.0)
        view = scene.fit_camera(meshes[g], w, h, heading=heading)
        params = plugin.Renderer.make_params(view, grooms[g], meshes[g], dom_size, dirs)
        out.append((g, w, h, view, params, BATCH_GROOMS[g][3]))
    return BatchScene(meshes, grooms, out, points, dirs)
