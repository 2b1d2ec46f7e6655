"""The offline baker over the C ABI (hm_bake, src/bake.rs:89-213) and Cem
Yuksel's HAIR strand format (src/hair_file.rs:50-130).

bake_strands(points, strands) -> HairMesh runs libhairmi.so's host baker;
read_hair(path) parses a .hair file into the same (points, strands) arrays
HairStrands holds.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _native as N
from .groom import HairMesh


def bake_strands(points: np.ndarray, strands: np.ndarray) -> HairMesh:
    """points f32[P, 3]; strands u32[S, 2] (first point, point count)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    st = np.ascontiguousarray(strands, np.uint32).reshape(-1, 2)
    out = C.c_void_p()
    N.check(N.lib().hm_bake(N.ptr(pts) if len(pts) else None, len(pts), N.ptr(st) if len(st) else None, len(st),
                            C.byref(out)), None, "hm_bake")
    try:
        d = N.HmGroomDesc()
        N.check(N.lib().hm_bake_get(out, C.byref(d)), None, "hm_bake_get")
        nb, nl, ns = d.bundle_count, d.layer_count, d.style_count

        def arr(ptr, count, width, dtype):
            if count == 0:
                return np.zeros((0, width), dtype)
            buf = (C.c_char * (count * width * 4)).from_address(ptr)
            return np.frombuffer(buf, dtype).reshape(count, width).copy()

        bundles = arr(d.bundles, nb, 4, np.uint32)
        layers = arr(d.layers, nl, 8, np.float32)
        style = arr(d.style, ns, 4, np.float32)
        return HairMesh(bundles, layers, style, np.array(d.bounds_min, np.float32),
                        np.array(d.bounds_max, np.float32))
    finally:
        N.lib().hm_bake_free(out)


def read_hair(path: str):
    """(points f32[P, 3], strands u32[S, 2], default colour) of a .hair file.
    The optional thickness / transparency / colour arrays are not read: the
    groom component carries the appearance (hair_file.rs:90-100)."""
    data = open(path, "rb").read()
    if len(data) < 128 or data[:4] != b"HAIR":
        raise ValueError(f"{path}: not a HAIR file")
    n_strands, n_points, flags, default_segments = struct.unpack_from("<4I", data, 4)
    color = struct.unpack_from("<3f", data, 28)
    if not flags & 2:
        raise ValueError(f"{path}: no point array")
    cursor = 128
    if flags & 1:
        segments = np.frombuffer(data, np.uint16, n_strands, cursor).astype(np.uint32)
        cursor += 2 * n_strands
    else:
        segments = np.full(n_strands, default_segments, np.uint32)
    if cursor + 12 * n_points > len(data):
        raise ValueError(f"{path}: truncated")
    points = np.frombuffer(data, np.float32, 3 * n_points, cursor).reshape(-1, 3).copy()
    counts = segments + 1
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    if n_strands and int(starts[-1] + counts[-1]) > n_points:
        raise ValueError(f"{path}: truncated")
    return points, np.stack([starts.astype(np.uint32), counts], 1), color
