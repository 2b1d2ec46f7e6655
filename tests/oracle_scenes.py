"""Hand-built inputs whose raster results can be enumerated by hand.

`screen_params` maps world (x, y) straight to pixel coordinates: clip x/y are
affine in x/y, clip w = 1, so project_world gives screen = (x, y) exactly for
values representable in float. The camera looks down +z from z = -1, so the
view distance of a point is z + 1; clip z is a constant `clip_depth`.
`line_mesh` is one bundle whose 16 layers are points on a straight line (all
four corners equal, zero style): every generated strand is that line.
"""
from __future__ import annotations

import numpy as np

from bevy_hair_amd import _native as N
from bevy_hair_amd.groom import HairMesh


def screen_params(width: int, height: int, mode: int = 0, bundles: int = 1, reverse: int = 1,
                  near: float = 0.05, far: float = 10.0, clip_depth: float = 0.5, lod: int = 0,
                  dom_size: int = 0, center_diameter: float = 1.0, ao: int = 1) -> N.HmParams:
    p = N.HmParams()
    m = np.zeros((4, 4))  # rows
    m[0] = [2.0 / width, 0.0, 0.0, -1.0]
    m[1] = [0.0, -2.0 / height, 0.0, 1.0]
    m[2] = [0.0, 0.0, -0.01, clip_depth]
    m[3] = [0.0, 0.0, 0.0, 1.0]
    p.clip_from_world[:] = [float(x) for x in m.T.reshape(-1)]
    p.world_from_clip[:] = [float(x) for x in np.linalg.inv(m).T.reshape(-1)]
    p.world_from_model[:] = [float(x) for x in np.eye(4).T.reshape(-1)]
    p.light_clip_from_world[:] = p.clip_from_world[:]
    p.camera_pos[:] = [0.0, 0.0, -1.0, 1.0]
    p.camera_forward[:] = [0.0, 0.0, 1.0, 0.0]
    p.screen[:] = [float(width), float(height), near, far]
    p.appearance[:] = [0.2, 0.08, 3.0, center_diameter]
    p.filter_params[:] = [1.0, 0.2, 0.1, 0.05]
    p.albedo[:] = [0.3, 0.2, 0.1, 1.0]
    p.flags[:] = [lod, 1, ao, 1]
    p.pass_mode[:] = [mode, bundles, reverse, 1 if mode else 0]
    p.light_eye[:] = [0.0, 0.0, -1.0, 1.0]
    p.light_forward[:] = [0.0, 0.0, 1.0, 0.0]
    S = dom_size or width
    p.dom_info[:] = [float(S), float(S), near, far]
    return p


def line_mesh(a, b, strands: int = 1) -> HairMesh:
    """One bundle; layer k of the cage is a + (b - a) * k / 15 (all corners)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    layers = np.zeros((64, 8), np.float32)
    for k in range(16):
        pnt = a + (b - a) * (k / 15.0)
        for c in range(4):
            layers[k * 4 + c, :3] = pnt
            layers[k * 4 + c, 3] = 1.0
    bundles = np.array([[0, 0, strands, 0]], np.uint32)
    style = np.zeros((1024, 4), np.float32)
    pos = layers[:, :3]
    return HairMesh(bundles, layers, style, pos.min(0), pos.max(0))


def refs_for(mesh: HairMesh, control_points: int, n_lod: int | None = None) -> np.ndarray:
    out = []
    for b, (_, _, n, _) in enumerate(mesh.bundles):
        k = n if n_lod is None else n_lod
        for s in range(k):
            out.append((b, s, k, control_points))
    return np.array(out, np.uint32).reshape(-1, 4)
