"""Reference-side helpers on a HairMesh (test infrastructure)."""
from __future__ import annotations

import numpy as np


def styled_position(mesh, b, uv, w):
    """float64 restatement of src/mesh.rs:100-175."""
    bundle = mesh.bundles[b]
    lay = mesh.layers.astype(np.float64)

    def corners(layer):
        o = int(bundle[0]) + layer * 4
        return lay[o:o + 4, :3]

    def bil(q):
        u, v = uv
        return q[0] * (1 - u) * (1 - v) + q[1] * u * (1 - v) + q[2] * (1 - u) * v + q[3] * u * v

    def tangent(layer):
        curr = bil(corners(layer))
        prev = bil(corners(max(layer - 1, 0)))
        nxt = bil(corners(min(layer + 1, 15)))
        if layer == 0:
            d, s = nxt - curr, np.linalg.norm(nxt - curr)
        elif layer + 1 >= 16:
            d, s = curr - prev, np.linalg.norm(curr - prev)
        else:
            d, s = nxt - prev, np.linalg.norm(nxt - prev) * 0.5
        n = np.linalg.norm(d)
        d = d / n if n > 0 else np.array([0.0, 1.0, 0.0])
        return d * max(s, 1e-4)

    lf = min(max(w, 0.0), 1.0) * 15.0
    i0 = min(int(np.floor(lf)), 15)
    i1 = min(i0 + 1, 15)
    t = lf - i0
    p0, p1 = bil(corners(i0)), bil(corners(i1))
    m0, m1 = tangent(i0), tangent(i1)
    t2, t3 = t * t, t * t * t
    base = p0 * (2 * t3 - 3 * t2 + 1) + m0 * (t3 - 2 * t2 + t) + p1 * (-2 * t3 + 3 * t2) + m1 * (t3 - t2)
    st = mesh.style.astype(np.float64)
    so = int(bundle[1])
    u, v, ww = min(max(uv[0], 0), 1) * 7, min(max(uv[1], 0), 1) * 7, min(max(w, 0), 1) * 15
    u0, v0, w0 = int(np.floor(u)), int(np.floor(v)), int(np.floor(ww))
    u1, v1, w1 = min(u0 + 1, 7), min(v0 + 1, 7), min(w0 + 1, 15)
    tu, tv, tw = u - u0, v - v0, ww - w0

    def at(a, b_, c):
        return st[so + (c * 8 + b_) * 8 + a, :3]

    def lerp(a, b_, t_):
        return a + (b_ - a) * t_

    c00 = lerp(at(u0, v0, w0), at(u1, v0, w0), tu)
    c01 = lerp(at(u0, v1, w0), at(u1, v1, w0), tu)
    c10 = lerp(at(u0, v0, w1), at(u1, v0, w1), tu)
    c11 = lerp(at(u0, v1, w1), at(u1, v1, w1), tu)
    return base + lerp(lerp(c00, c01, tv), lerp(c10, c11, tv), tw)


def styled_positions(mesh, bundle: np.ndarray, uv: np.ndarray, w: np.ndarray) -> np.ndarray:
    """styled_position for N queries at once (float64): bundle[N], uv[N, 2], w[N]."""
    bundle = np.asarray(bundle, np.int64)
    uv = np.asarray(uv, np.float64)
    w = np.asarray(w, np.float64)
    if getattr(mesh, "_f64", None) is None:  # float64 copies, made once per mesh
        mesh._f64 = (mesh.layers.astype(np.float64), mesh.style.astype(np.float64))
    lay, st = mesh._f64
    off = mesh.bundles[bundle, 0].astype(np.int64)
    u, v = uv[:, 0:1], uv[:, 1:2]

    def bil(layer):
        base = off + layer * 4
        q = [lay[base + c, :3] for c in range(4)]
        return q[0] * (1 - u) * (1 - v) + q[1] * u * (1 - v) + q[2] * (1 - u) * v + q[3] * u * v

    def tangent(layer):
        curr, prev, nxt = bil(layer), bil(np.maximum(layer - 1, 0)), bil(np.minimum(layer + 1, 15))
        first, last = (layer == 0)[:, None], (layer + 1 >= 16)[:, None]
        d = np.where(first, nxt - curr, np.where(last, curr - prev, nxt - prev))
        s = np.linalg.norm(d, axis=1, keepdims=True) * np.where(first | last, 1.0, 0.5)
        n = np.linalg.norm(d, axis=1, keepdims=True)
        d = np.where(n > 0, d / np.where(n > 0, n, 1.0), np.array([0.0, 1.0, 0.0]))
        return d * np.maximum(s, 1e-4)

    lf = np.clip(w, 0.0, 1.0) * 15.0
    i0 = np.minimum(np.floor(lf).astype(np.int64), 15)
    i1 = np.minimum(i0 + 1, 15)
    t = (lf - i0)[:, None]
    p0, p1, m0, m1 = bil(i0), bil(i1), tangent(i0), tangent(i1)
    t2, t3 = t * t, t * t * t
    base = p0 * (2 * t3 - 3 * t2 + 1) + m0 * (t3 - 2 * t2 + t) + p1 * (-2 * t3 + 3 * t2) + m1 * (t3 - t2)
    so = mesh.bundles[bundle, 1].astype(np.int64)
    su, sv, sw = np.clip(uv[:, 0], 0, 1) * 7, np.clip(uv[:, 1], 0, 1) * 7, np.clip(w, 0, 1) * 15
    u0, v0, w0 = np.floor(su).astype(np.int64), np.floor(sv).astype(np.int64), np.floor(sw).astype(np.int64)
    u1, v1, w1 = np.minimum(u0 + 1, 7), np.minimum(v0 + 1, 7), np.minimum(w0 + 1, 15)
    tu, tv, tw = (su - u0)[:, None], (sv - v0)[:, None], (sw - w0)[:, None]

    def at(a, b_, c):
        return st[so + (c * 8 + b_) * 8 + a, :3]

    def lerp(a, b_, t_):
        return a + (b_ - a) * t_

    c00 = lerp(at(u0, v0, w0), at(u1, v0, w0), tu)
    c01 = lerp(at(u0, v1, w0), at(u1, v1, w0), tu)
    c10 = lerp(at(u0, v0, w1), at(u1, v0, w1), tu)
    c11 = lerp(at(u0, v1, w1), at(u1, v1, w1), tu)
    return base + lerp(lerp(c00, c01, tv), lerp(c10, c11, tv), tw)
