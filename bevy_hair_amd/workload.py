"""A ready-to-render workload: seeded groom, fitted camera, lights and HairParams."""
from __future__ import annotations

import dataclasses

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
