"""The baker (hm_bake, src/bake.rs:89-213) against the reference's own tests
(bake.rs:1391-1527) on their synthetic fixtures, and on the reference's
example grooms where the reference checkout is present (read in place).

The stray distance is the reference's max_styled_stray (bake.rs:1554-1600):
styled positions (mesh.rs::styled_position, the float64 restatement in
tests/mesh_ref.py) at a 5 x 5 uv grid and 31 steps along each bundle,
measured to the nearest source segment.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from bevy_hair_amd.bake import bake_strands, read_hair
from tests.mesh_ref import styled_position

LAYERS = 16


def straight_bundle():
    """bake.rs:1529-1552: three strands at each corner of a 2 x 2 square, 8 points up +y."""
    pts, st = [], []
    corners = [(-1.0, 0.0, -1.0), (1.0, 0.0, -1.0), (-1.0, 0.0, 1.0), (1.0, 0.0, 1.0)]
    for _ in range(3):
        for c in corners:
            st.append((len(pts), 8))
            pts += [(c[0], c[1] + layer, c[2]) for layer in range(8)]
    return np.array(pts, np.float32), np.array(st, np.uint32)


def diverging_locks():
    """bake.rs:1434-1450: two 16-strand locks from the origin toward x = -30 and x = +30."""
    pts, st = [], []
    for side in (-1.0, 1.0):
        for i in range(16):
            tip = np.array([side * 30.0, (i - 7.5) * 0.15, -40.0])
            st.append((len(pts), 16))
            pts += [tuple(tip * (layer / 15.0)) for layer in range(16)]
    return np.array(pts, np.float32), np.array(st, np.uint32)


def bent_lock():
    """bake.rs:1466-1482: a 4 x 4 lock bending a quarter turn over 16 layers."""
    pts, st = [], []
    for i in range(4):
        for j in range(4):
            root = np.array([i * 0.4, j * 0.4, 0.0])
            st.append((len(pts), 16))
            for layer in range(16):
                a = layer / 15.0 * np.pi / 2
                pts.append(tuple(root + [20.0 * (1.0 - np.cos(a)), 0.0, 20.0 * np.sin(a)]))
    return np.array(pts, np.float32), np.array(st, np.uint32)


def segment_distances(pos: np.ndarray, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance from each of pos[K, 3] to the nearest of the segments a[M] -> b[M]."""
    ab = b - a
    den = np.maximum((ab * ab).sum(1), 1e-8)
    out = np.empty(len(pos))
    for k0 in range(0, len(pos), 256):
        p = pos[k0:k0 + 256, None, :]
        t = np.clip(((p - a[None]) * ab[None]).sum(2) / den[None], 0.0, 1.0)
        d = p - (a[None] + ab[None] * t[..., None])
        out[k0:k0 + 256] = np.sqrt((d * d).sum(2)).min(1)
    return out


def max_styled_stray(mesh, points, strands) -> float:
    starts = strands[:, 0].astype(np.int64)
    counts = strands[:, 1].astype(np.int64)
    a_idx = np.concatenate([np.arange(s, s + c - 1) for s, c in zip(starts, counts) if c >= 2])
    a, b = points[a_idx].astype(np.float64), points[a_idx + 1].astype(np.float64)
    uvs = [0.0, 0.25, 0.5, 0.75, 1.0]
    pos = np.array([styled_position(mesh, bi, (u, v), step / (LAYERS * 2.0 - 2.0))
                    for bi in range(mesh.bundle_count) for v in uvs for u in uvs for step in range(LAYERS * 2 - 1)])
    return float(segment_distances(pos, a, b).max())


def test_straight_cage_residuals_are_near_zero():
    mesh = bake_strands(*straight_bundle())
    assert mesh.bundle_count == 1
    assert np.isfinite(mesh.layers[:, :7]).all()
    assert np.abs(mesh.style[:, :3]).sum(1).max() < 1e-2 * 1.0  # |residual| < 1e-2 (L1 bounds L2)
    assert int(mesh.bundles[0, 2]) == 12
    mesh.validate()


def test_isolated_bundle_ao_does_not_darken_toward_the_root():
    mesh = bake_strands(*straight_bundle())
    ao = mesh.layers[:, 3].reshape(LAYERS, 4).mean(1)
    assert ao[0] + 0.15 >= ao[-1], (ao[0], ao[-1])
    assert ((ao >= 0.0) & (ao <= 1.0)).all()


def test_diverging_strands_are_not_bridged():
    pts, st = diverging_locks()
    mesh = bake_strands(pts, st)
    assert mesh.bundle_count > 1, "the two locks should not share one quad"
    assert int(mesh.bundles[:, 2].sum()) == 32
    stray = max_styled_stray(mesh, pts, st)
    assert stray < 8.0, stray


def test_bent_bundle_keeps_its_cross_section():
    pts, st = bent_lock()
    mesh = bake_strands(pts, st)
    stray = max_styled_stray(mesh, pts, st)
    assert stray < 8.0, stray


def test_empty_and_degenerate_inputs():
    mesh = bake_strands(np.zeros((0, 3), np.float32), np.zeros((0, 2), np.uint32))
    assert mesh.bundle_count == 0
    # one-point strands and zero-length strands are dropped (bake.rs:683-714)
    pts = np.zeros((3, 3), np.float32)
    mesh = bake_strands(pts, np.array([[0, 1], [1, 2]], np.uint32))
    assert mesh.bundle_count == 0


ASSETS = "/root/reference/assets/hair"


def resample16(points: np.ndarray, start: int, count: int) -> np.ndarray:
    """16 arc-length samples of one strand (bake.rs:683-714), float64."""
    p = points[start:start + count].astype(np.float64)
    cum = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(p, axis=0), axis=1))])
    target = cum[-1] * np.arange(16) / 15.0
    out = np.empty((16, 3))
    for d in range(3):
        out[:, d] = np.interp(target, cum, p[:, d])
    return out


def quad_uv_and_plane(q: np.ndarray, pts: np.ndarray):
    """mesh.rs quad_uv and quad_plane_distance for quads q[N, 4, 3] and points pts[N, 3]."""
    eu, ev, rel = q[:, 1] - q[:, 0], q[:, 2] - q[:, 0], pts - q[:, 0]
    n = np.cross(eu, ev)
    den = (n * n).sum(1)
    u = (np.cross(rel, ev) * n).sum(1) / den
    v = (np.cross(eu, rel) * n).sum(1) / den
    return u, v, (rel * n).sum(1) / np.sqrt(den)


def inside(u, v):
    return (u >= -1e-3) & (v >= -1e-3) & (u <= 1 + 1e-3) & (v <= 1 + 1e-3)


@pytest.mark.skipif(not os.path.isdir(ASSETS), reason="the reference checkout's assets are not present")
@pytest.mark.parametrize("name", ["wStraight", "wWavy", "wCurly"])
def test_asset_grooms_stay_on_the_source_strands(name):
    """bake.rs:1496-1527 on the example grooms (3D-grid clustering): no strand
    dropped; every root inside some bundle's root quad (uv in the unit square,
    within 8 of its plane); styled positions (5 x 5 uv x 31 steps) of 1500
    seeded bundles within 8 of the source strands; 1000 seeded strands
    reconstructed within 6 by their best containing bundle. Distances to the
    source go to points sampled every <= 1 unit along its segments, never
    nearer than the segments themselves, so the bound is conservative."""
    from scipy.spatial import cKDTree

    from tests.mesh_ref import styled_positions

    pts, st, _ = read_hair(os.path.join(ASSETS, f"{name}.hair"))
    mesh = bake_strands(pts, st)
    usable = st[:, 1] >= 2
    assert int(mesh.bundles[:, 2].sum()) == int(usable.sum()), "strands dropped"
    nb = mesh.bundle_count
    lay = mesh.layers.astype(np.float64)
    root_quads = np.stack([lay[mesh.bundles[:, 0].astype(np.int64) + c, :3] for c in range(4)], 1)  # [nb, 4, 3]
    ids = np.nonzero(usable)[0]
    roots = pts[st[ids, 0]].astype(np.float64)

    def contains(r: np.ndarray, bundles: np.ndarray):
        """[len(r), len(bundles)] containment of roots r in the bundles' root quads, and the uv."""
        q = np.broadcast_to(root_quads[bundles][None], (len(r), len(bundles), 4, 3)).reshape(-1, 4, 3)
        u, v, h = quad_uv_and_plane(q, np.repeat(r, len(bundles), 0))
        ok = inside(u, v) & (np.abs(h) < 8.0)
        return ok.reshape(len(r), -1), u.reshape(len(r), -1), v.reshape(len(r), -1)

    # A quad that contains a root (uv in the unit square, 8 from its plane; the
    # slack of 1e-3 in uv included) has its centre within its half diagonal
    # times 1.002 plus 8 of the root: a ball query finds every candidate.
    centres = root_quads.mean(1)
    half_diag = 0.5 * np.maximum(np.linalg.norm(root_quads[:, 3] - root_quads[:, 0], axis=1),
                                 np.linalg.norm(root_quads[:, 2] - root_quads[:, 1], axis=1))
    reach = float(half_diag.max()) * 1.002 + 8.0 + 1e-3
    centre_tree = cKDTree(centres)

    def candidates(r: np.ndarray) -> np.ndarray:
        return np.asarray(centre_tree.query_ball_point(r, reach), np.int64)

    _, cand = centre_tree.query(roots, k=min(16, nb))
    R, K = cand.shape
    u, v, h = quad_uv_and_plane(root_quads[cand].reshape(R * K, 4, 3), np.repeat(roots, K, 0))
    covered = (inside(u, v) & (np.abs(h) < 8.0)).reshape(R, K).any(1)
    for i in np.nonzero(~covered)[0]:
        covered[i] = contains(roots[i:i + 1], candidates(roots[i]))[0].any()
    assert covered.all(), f"{(~covered).sum()} roots outside every root quad"
    # stray of 1500 seeded bundles
    rng = np.random.default_rng(7)
    a_idx = np.concatenate([np.arange(s0, s0 + c - 1) for s0, c in st[usable]])
    a, b = pts[a_idx].astype(np.float64), pts[a_idx + 1].astype(np.float64)
    k = np.maximum(1, np.ceil(np.linalg.norm(b - a, axis=1))).astype(np.int64)
    rep = np.repeat(np.arange(len(a)), k)
    t = (np.arange(len(rep)) - np.repeat(np.cumsum(k) - k, k)) / np.repeat(k, k)
    tree = cKDTree(np.concatenate([a[rep] + (b[rep] - a[rep]) * t[:, None], b]))
    grid = np.array([0.0, 0.25, 0.5, 0.75, 1.0])
    U, V, Wq = np.meshgrid(grid, grid, np.arange(31) / 30.0, indexing="ij")
    uvw = np.stack([U.ravel(), V.ravel(), Wq.ravel()], 1)
    sel = rng.choice(nb, min(1500, nb), replace=False)
    bb = np.repeat(sel, len(uvw))
    qs = np.tile(uvw, (len(sel), 1))
    stray = float(tree.query(styled_positions(mesh, bb, qs[:, :2], qs[:, 2]))[0].max())
    assert stray < 8.0, f"{name}: styled cage leaves the source groom by {stray:.2f}"
    # reconstruction of 1000 seeded strands by their best containing bundle
    pick = rng.choice(len(ids), min(1000, len(ids)), replace=False)
    worst = 0.0
    for i in pick:
        cb = candidates(roots[i])
        ok, uu, vv = contains(roots[i:i + 1], cb)
        bundles, uu, vv = cb[ok[0]], uu[0, ok[0]], vv[0, ok[0]]
        s_i = ids[i]
        samples = resample16(pts, int(st[s_i, 0]), int(st[s_i, 1]))
        n = len(bundles)
        assert n > 0
        pos = styled_positions(mesh, np.repeat(bundles, 16), np.repeat(np.stack([uu, vv], 1), 16, 0),
                               np.tile(np.arange(16) / 15.0, n)).reshape(n, 16, 3)
        worst = max(worst, float(np.linalg.norm(pos - samples[None], axis=2).max(1).min()))
    print(f"{name}: {nb} bundles, stray {stray:.2f}, reconstruction {worst:.2f}")
    assert worst < 6.0, f"{name}: reconstruction misses a strand by {worst:.2f}"
