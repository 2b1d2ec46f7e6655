"""HairMesh (mesh.rs:31-41), seeded synthetic grooms and the HAIRBAK1 groom cache.

A HairMesh is what the rasterizer consumes: bundles of 16 layers x 4 cage
corners (c00, c10, c01, c11) plus an 8x8x16 style volume of residuals per
bundle. The reference produces it with an offline baker (src/bake.rs); the
synthetic generator here follows the baker's 3D-grid path in outline only —
roots clustered into ~180-strand bundles, each strand resampled at 16 arc-length
layers (bake.rs:683-714), a PCA root frame with transported per-layer axes and
tight bounds (bake.rs:716-826), corner tangents by finite differences
(bake.rs:166-184), the mean style residual splatted bilinearly
(bake.rs:1099-1123), AO = 1. It is input preparation, not part of the parity
path: both the HIP kernels and the CPU restatement consume the same arrays.
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np

LAYER_COUNT = 16
STYLE_U, STYLE_V, STYLE_W = 8, 8, 16
STYLE_TEXELS = STYLE_U * STYLE_V * STYLE_W
TARGET_STRANDS_PER_BUNDLE = 180  # bake.rs:18


@dataclasses.dataclass
class HairMesh:
    """bundles u32[B,4] (layer_offset, style_offset, strand_count, 0);
    layers f32[B*64, 8] (pos.xyz, ao, tangent.xyz, 0); style f32[B*1024, 4]."""

    bundles: np.ndarray
    layers: np.ndarray
    style: np.ndarray
    bounds_min: np.ndarray
    bounds_max: np.ndarray

    @property
    def bundle_count(self) -> int:
        return int(self.bundles.shape[0])

    @property
    def strand_count(self) -> int:
        return int(self.bundles[:, 2].sum()) if self.bundle_count else 0

    def validate(self) -> None:
        assert self.bundles.dtype == np.uint32 and self.bundles.ndim == 2 and self.bundles.shape[1] == 4
        assert self.layers.dtype == np.float32 and self.layers.shape[1] == 8
        assert self.style.dtype == np.float32 and self.style.shape[1] == 4
        for b in self.bundles:
            assert b[0] + 64 <= len(self.layers) and b[1] + STYLE_TEXELS <= len(self.style)


# ---- strands ---------------------------------------------------------------------


def _cap_roots(rng: np.random.Generator, n: int, max_polar_deg: float = 70.0) -> np.ndarray:
    """Uniform by area on the unit-sphere cap around +Y with polar angle < max."""
    cmin = np.cos(np.radians(max_polar_deg))
    cos_t = rng.uniform(cmin, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t**2))
    return np.stack([sin_t * np.cos(phi), cos_t, sin_t * np.sin(phi)], axis=1)


def _frame(normals: np.ndarray):
    """Per-strand orthonormal frame (n, a, b) with `a` horizontal where possible."""
    up = np.array([0.0, 1.0, 0.0])
    a = np.cross(normals, up)
    la = np.linalg.norm(a, axis=1, keepdims=True)
    a = np.where(la > 1e-6, a / np.maximum(la, 1e-12), np.array([1.0, 0.0, 0.0]))
    b = np.cross(normals, a)
    return a, b


def synth_strands(kind: str, n: int, seed: int, points: int = 48) -> np.ndarray:
    """Seeded synthetic strands, f64[n, points, 3] on a unit-sphere scalp.

    straight: length 0.3 +-20%, gravity droop 0.5 (config C1).
    wavy:     length 0.35 +-25%, lateral sine, amplitude 0.01, period 0.05, random phase (C2).
    curly:    helix radius 0.008 +-30%, pitch 0.02 +-20% around a drooping axis (C3).
    """
    rng = np.random.default_rng(seed)
    roots = _cap_roots(rng, n)
    normals = roots / np.linalg.norm(roots, axis=1, keepdims=True)
    if kind == "straight":
        length = 0.3 * rng.uniform(0.8, 1.2, n)
    elif kind == "wavy":
        length = 0.35 * rng.uniform(0.75, 1.25, n)
    elif kind == "curly":
        length = 0.35 * rng.uniform(0.8, 1.2, n)
    else:
        raise ValueError(kind)
    droop = 0.5
    s = np.linspace(0.0, 1.0, points)  # normalized arc parameter
    gravity = np.array([0.0, -1.0, 0.0])
    # Axis curve: leaves along the normal, bends toward gravity.
    axis = (normals[:, None, :] * s[None, :, None] + droop * 0.5 * (s**2)[None, :, None] * gravity) * length[:, None, None]
    pts = roots[:, None, :] + axis
    a, b = _frame(normals)
    arc = s[None, :] * length[:, None]
    if kind == "wavy":
        phase = rng.uniform(0.0, 2.0 * np.pi, n)
        off = 0.01 * np.sin(2.0 * np.pi * arc / 0.05 + phase[:, None])
        pts = pts + off[:, :, None] * a[:, None, :]
    elif kind == "curly":
        radius = 0.008 * rng.uniform(0.7, 1.3, n)
        pitch = 0.02 * rng.uniform(0.8, 1.2, n)
        phase = rng.uniform(0.0, 2.0 * np.pi, n)
        ang = 2.0 * np.pi * arc / pitch[:, None] + phase[:, None]
        # the helix grows in from the root over the first 5% of the strand
        ramp = np.clip(s / 0.05, 0.0, 1.0)[None, :]
        pts = pts + (radius[:, None] * ramp * np.cos(ang))[:, :, None] * a[:, None, :] \
                  + (radius[:, None] * ramp * np.sin(ang))[:, :, None] * b[:, None, :]
    return pts


def resample(strands: np.ndarray, layers: int = LAYER_COUNT) -> np.ndarray:
    """Arc-length resampling at `layers` samples per strand (bake.rs:683-714)."""
    seg = np.linalg.norm(np.diff(strands, axis=1), axis=2)
    cum = np.concatenate([np.zeros((len(strands), 1)), np.cumsum(seg, axis=1)], axis=1)
    length = cum[:, -1:]
    target = length * (np.arange(layers) / (layers - 1))[None, :]
    out = np.empty((len(strands), layers, 3))
    for i in range(len(strands)):
        for d in range(3):
            out[i, :, d] = np.interp(target[i], cum[i], strands[i, :, d])
    return out


# ---- clustering and cages ---------------------------------------------------------


def _fibonacci_cap(k: int, max_polar_deg: float = 70.0) -> np.ndarray:
    cmin = np.cos(np.radians(max_polar_deg))
    i = np.arange(k) + 0.5
    cos_t = 1.0 - (1.0 - cmin) * i / k
    phi = np.pi * (3.0 - np.sqrt(5.0)) * i
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t**2))
    return np.stack([sin_t * np.cos(phi), cos_t, sin_t * np.sin(phi)], axis=1)


def cluster_roots(roots: np.ndarray, target: int = TARGET_STRANDS_PER_BUNDLE) -> list[np.ndarray]:
    """Nearest-centre clustering on a spiral of ~N/target centres (spatially ordered)."""
    k = max(1, int(np.ceil(len(roots) / target)))
    centres = _fibonacci_cap(k)
    assign = np.empty(len(roots), dtype=np.int64)
    step = 65536
    for lo in range(0, len(roots), step):
        r = roots[lo:lo + step]
        d = ((r[:, None, :] - centres[None, :, :]) ** 2).sum(axis=2)
        assign[lo:lo + step] = d.argmin(axis=1)
    order = np.argsort(assign, kind="stable")
    counts = np.bincount(assign, minlength=k)
    groups = np.split(order, np.cumsum(counts)[:-1])
    return [g for g in groups if len(g) > 0]


def _tight(points, mean, au, av):
    d = points - mean
    lu, lv = d @ au, d @ av
    mnu, mxu, mnv, mxv = lu.min(), lu.max(), lv.min(), lv.max()
    pu = max(mxu - mnu, 1e-3) * 1e-3 + 1e-4
    pv = max(mxv - mnv, 1e-3) * 1e-3 + 1e-4
    mnu, mxu, mnv, mxv = mnu - pu, mxu + pu, mnv - pv, mxv + pv
    if mxu - mnu < 1e-2:
        mid = 0.5 * (mxu + mnu)
        mnu, mxu = mid - 5e-3, mid + 5e-3
    if mxv - mnv < 1e-2:
        mid = 0.5 * (mxv + mnv)
        mnv, mxv = mid - 5e-3, mid + 5e-3
    return mnu, mxu, mnv, mxv


def _transport(pu, pv, normal):
    up_ = pu - normal * (pu @ normal)
    vp_ = pv - normal * (pv @ normal)
    if up_ @ up_ >= vp_ @ vp_ and up_ @ up_ > 1e-8:
        au = up_ / np.linalg.norm(up_)
        av = np.cross(normal, au)
        if av @ pv < 0:
            av = -av
        return au, av / np.linalg.norm(av)
    if vp_ @ vp_ > 1e-8:
        av = vp_ / np.linalg.norm(vp_)
        au = np.cross(av, normal)
        if au @ pu < 0:
            au = -au
        return au / np.linalg.norm(au), av
    au = np.cross(normal, [0.0, 1.0, 0.0])
    au = au / max(np.linalg.norm(au), 1e-12)
    return au, np.cross(normal, au)


def bake_bundles(samples: np.ndarray, groups: list[np.ndarray]) -> HairMesh:
    """Cages for each group of resampled strands (samples f64[N, 16, 3])."""
    bundles, layers, style = [], [], []
    for g in groups:
        s = samples[g]  # [m, 16, 3]
        roots = s[:, 0, :]
        mean = roots.mean(axis=0)
        d = roots - mean
        cov = d.T @ d
        w, v = np.linalg.eigh(cov)
        au = v[:, 2]
        av = v[:, 1]
        av = av - au * (av @ au)
        av = av / max(np.linalg.norm(av), 1e-12)
        quads = np.empty((LAYER_COUNT, 4, 3))
        frames = []
        for L in range(LAYER_COUNT):
            pts = s[:, L, :]
            m = pts.mean(axis=0)
            if L > 0:
                nxt = s[:, min(L + 1, LAYER_COUNT - 1), :]
                prv = s[:, L - 1, :]
                t = (nxt - prv).sum(axis=0)
                nt = np.linalg.norm(t)
                normal = t / nt if nt > 1e-12 else np.array([0.0, 1.0, 0.0])
                au, av = _transport(au, av, normal)
            mnu, mxu, mnv, mxv = _tight(pts, m, au, av)
            frames.append((m, au, av, mnu, mxu, mnv, mxv))
            quads[L] = [m + au * mnu + av * mnv, m + au * mxu + av * mnv,
                        m + au * mnu + av * mxv, m + au * mxu + av * mxv]
        # style: mean residual of members from the bilinear cage at their root uv
        m0, au0, av0, mnu0, mxu0, mnv0, mxv0 = frames[0]
        ru = ((roots - m0) @ au0 - mnu0) / (mxu0 - mnu0)
        rv = ((roots - m0) @ av0 - mnv0) / (mxv0 - mnv0)
        ru, rv = np.clip(ru, 0, 1), np.clip(rv, 0, 1)
        ssum = np.zeros((STYLE_TEXELS, 3))
        swt = np.zeros(STYLE_TEXELS)
        uu, vv = ru * (STYLE_U - 1), rv * (STYLE_V - 1)
        u0, v0 = np.floor(uu).astype(int), np.floor(vv).astype(int)
        u1, v1 = np.minimum(u0 + 1, STYLE_U - 1), np.minimum(v0 + 1, STYLE_V - 1)
        tu, tv = uu - u0, vv - v0
        for L in range(LAYER_COUNT):
            q = quads[L]
            bil = (q[0] * ((1 - ru) * (1 - rv))[:, None] + q[1] * (ru * (1 - rv))[:, None]
                   + q[2] * ((1 - ru) * rv)[:, None] + q[3] * (ru * rv)[:, None])
            res = s[:, L, :] - bil
            for iu, iv, alpha in ((u0, v0, (1 - tu) * (1 - tv)), (u1, v0, tu * (1 - tv)),
                                  (u0, v1, (1 - tu) * tv), (u1, v1, tu * tv)):
                idx = (L * STYLE_V + iv) * STYLE_U + iu
                np.add.at(ssum, idx, res * alpha[:, None])
                np.add.at(swt, idx, alpha)
        st = np.zeros((STYLE_TEXELS, 4))
        ok = swt > 0
        st[ok, :3] = ssum[ok] / swt[ok, None]
        # corners with finite-difference tangents along the layers (bake.rs:166-184)
        lay = np.zeros((LAYER_COUNT * 4, 8))
        for L in range(LAYER_COUNT):
            prev, nxt = quads[max(L - 1, 0)], quads[min(L + 1, LAYER_COUNT - 1)]
            if L == 0:
                tan = nxt - quads[L]
            elif L == LAYER_COUNT - 1:
                tan = quads[L] - prev
            else:
                tan = (nxt - prev) * 0.5
            for c in range(4):
                lay[L * 4 + c, :3] = quads[L, c]
                lay[L * 4 + c, 3] = 1.0
                lay[L * 4 + c, 4:7] = tan[c]
        bundles.append((len(layers) * 64, len(style) * STYLE_TEXELS, len(g), 0))
        layers.append(lay)
        style.append(st)
    lay = np.concatenate(layers).astype(np.float32)
    sty = np.concatenate(style).astype(np.float32)
    pos = lay[:, :3]
    return HairMesh(np.array(bundles, dtype=np.uint32), np.ascontiguousarray(lay), np.ascontiguousarray(sty),
                    pos.min(axis=0).astype(np.float32), pos.max(axis=0).astype(np.float32))


def bake(strands: np.ndarray, target: int = TARGET_STRANDS_PER_BUNDLE) -> HairMesh:
    samples = resample(strands)
    return bake_bundles(samples, cluster_roots(samples[:, 0, :], target))


# Cached synthetic grooms (generation of the large ones takes seconds).
_CACHE: dict = {}


def synthetic_groom(kind: str, n: int, seed: int) -> HairMesh:
    key = (kind, n, seed)
    if key not in _CACHE:
        _CACHE[key] = bake(synth_strands(kind, n, seed))
    return _CACHE[key]


# ---- HAIRBAK1 groom cache (examples/bake.rs:91-241) --------------------------------

CACHE_MAGIC = b"HAIRBAK1"
# FNV-1a of the reference baker's source text ^ (16 << 32) ^ 1024 (bake.rs:34-47).
BAKE_FINGERPRINT = 0x1A96CD5F1CA06552


@dataclasses.dataclass
class SourceStamp:
    len: int = 0
    secs: int = 0
    nanos: int = 0
    scalp_len: int = 0
    scalp_secs: int = 0
    scalp_nanos: int = 0


def encode_cache(stamp: SourceStamp, mesh: HairMesh, fingerprint: int = BAKE_FINGERPRINT) -> bytes:
    """Byte layout of examples/bake.rs:153-186 (little-endian)."""
    out = bytearray(CACHE_MAGIC)
    out += struct.pack("<QQQIQQI", fingerprint, stamp.len, stamp.secs, stamp.nanos, stamp.scalp_len,
                       stamp.scalp_secs, stamp.scalp_nanos)
    out += struct.pack("<III", mesh.bundle_count, len(mesh.layers), len(mesh.style))
    out += struct.pack("<3f3f", *mesh.bounds_min.tolist(), *mesh.bounds_max.tolist())
    out += np.ascontiguousarray(mesh.bundles[:, :3], dtype="<u4").tobytes()
    corners = np.concatenate([mesh.layers[:, 0:3], mesh.layers[:, 4:7], mesh.layers[:, 3:4]], axis=1)
    out += np.ascontiguousarray(corners, dtype="<f4").tobytes()
    out += np.ascontiguousarray(mesh.style, dtype="<f4").tobytes()
    return bytes(out)


def decode_cache(data: bytes, stamp: SourceStamp, fingerprint: int = BAKE_FINGERPRINT) -> HairMesh | None:
    """examples/bake.rs:188-241: None on wrong magic, fingerprint, stamps or payload length."""
    head = 8 + struct.calcsize("<QQQIQQI") + 12 + 24
    if len(data) < head or data[:8] != CACHE_MAGIC:
        return None
    fp, ln, secs, nanos, sl, ss, sn = struct.unpack_from("<QQQIQQI", data, 8)
    if (fp, ln, secs, nanos, sl, ss, sn) != (fingerprint, stamp.len, stamp.secs, stamp.nanos, stamp.scalp_len,
                                             stamp.scalp_secs, stamp.scalp_nanos):
        return None
    off = 8 + struct.calcsize("<QQQIQQI")
    nb, nc, ns = struct.unpack_from("<III", data, off)
    off += 12
    bmin = np.array(struct.unpack_from("<3f", data, off), dtype=np.float32)
    bmax = np.array(struct.unpack_from("<3f", data, off + 12), dtype=np.float32)
    off += 24
    if len(data) - off != nb * 12 + nc * 28 + ns * 16:
        return None
    b = np.frombuffer(data, dtype="<u4", count=nb * 3, offset=off).reshape(nb, 3)
    off += nb * 12
    c = np.frombuffer(data, dtype="<f4", count=nc * 7, offset=off).reshape(nc, 7)
    off += nc * 28
    s = np.frombuffer(data, dtype="<f4", count=ns * 4, offset=off).reshape(ns, 4)
    bundles = np.zeros((nb, 4), dtype=np.uint32)
    bundles[:, :3] = b
    layers = np.zeros((nc, 8), dtype=np.float32)
    layers[:, 0:3], layers[:, 4:7], layers[:, 3] = c[:, 0:3], c[:, 3:6], c[:, 6]
    return HairMesh(bundles, layers, np.ascontiguousarray(s, dtype=np.float32), bmin, bmax)


def write_cache(path: str, stamp: SourceStamp, mesh: HairMesh) -> None:
    """Atomic write through a temporary file and rename (examples/bake.rs:136-151)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(encode_cache(stamp, mesh))
    os.replace(tmp, path)


def read_cache(path: str, stamp: SourceStamp) -> HairMesh | None:
    try:
        with open(path, "rb") as f:
            return decode_cache(f.read(), stamp)
    except OSError:
        return None
