"""Training of the learned shading term on the GPU with the library's own
kernels (hm_mlp_collect, hm_mlp_train: hand-written MFMA forward/backward and
Adam), and loading of trained weights.

Samples are per-pixel (features, targets) pairs of plain-method frames: the
c1 and c2 workloads rendered under random key-light directions (64 by
default), the targets being what the plain method computes for the terms
the network replaces. The saved .npz holds the float32 weights and biases,
the trained feature range (tiles outside it fall back to the plain method)
and the per-output loss scale.

    python -m bevy_hair_amd.mlp_train --out bevy_hair_amd/weights/mlp_key.npz
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import sys
import time

import numpy as np

from . import _native as N
from . import plugin
from .workload import make_scene

LOG_EPS = 1e-6  # MLP_LOG_EPS of csrc/hm_mlp.h
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "mlp_key.npz")


def key_directions(count: int, seed: int) -> np.ndarray:
    """Unit directions to the key light, uniform over the sphere."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(count, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def collect(r: plugin.Renderer, names, directions: np.ndarray, total: int, seed: int, variants: int = 4,
            log=print) -> int:
    """Renders every workload under every key direction with the plain method
    (`variants` frames per direction: the workload's own albedo, then random
    ones) and appends a seeded share of each frame's hit pixels to the GPU
    sample buffer, so that each workload contributes about total / len(names)
    samples. Returns the count."""
    r.mlp_reserve_samples(total)
    frames = [make_scene(name) for name in names]
    hits = []  # hit pixels of one frame of each workload
    for sc in frames:
        w = sc.w
        r.targets(w.width, w.height, w.dom_size)
        r.set_scene_depth(None, 0.0)
        r.set_lights(sc.points, sc.dirs)
        g = r.upload(sc.mesh)
        r.hair_pass(g, sc.params, r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction))
        c = r.gbuffer()[0]
        hits.append(int((c != np.uint64(0xFFFFFFFFFFFFFFFF)).sum()))
        g.free()
    rng = np.random.default_rng(seed)
    n = 0
    for sc, h in zip(frames, hits):
        w = sc.w
        keep = min(1.0, total / len(frames) / max(1, h * len(directions) * variants))
        r.targets(w.width, w.height, w.dom_size)
        r.set_scene_depth(None, 0.0)
        g = r.upload(sc.mesh)
        opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
        t0 = time.perf_counter()
        for k, d in enumerate(directions):
            dirs = [dict(sc.dirs[0], direction_to_light=[*d.tolist(), 0.0])] + list(sc.dirs[1:])
            r.set_lights(sc.points, dirs)
            for v in range(variants):
                # variant 0 keeps the workload's own albedo (c1: the HairGroom
                # default; c2: its drawn one); the others draw one log-uniformly
                # in [0.004, 0.6] per channel, which spans both
                groom = sc.groom
                if v:
                    a = np.exp(rng.uniform(np.log(0.004), np.log(0.6), 3))
                    groom = dataclasses.replace(sc.groom, albedo_linear=tuple(a.tolist()))
                params = plugin.Renderer.make_params(sc.view, groom, sc.mesh, w.dom_size, dirs)
                r.hair_pass(g, params, opts)
                n = r.mlp_collect(params[0], keep, seed * 100000 + k * 16 + v)
        log(f"[mlp] {w.name}: {len(directions)} key directions x {variants} albedos, keep {keep:.3f}, {n} samples so far "
            f"({time.perf_counter() - t0:.1f}s)")
        g.free()
    return n


def feature_range(x: np.ndarray, margin: float = 0.02):
    """Per-feature range of the training set (min / max, widened by `margin` of the span)."""
    lo = x.min(axis=0)
    hi = x.max(axis=0)
    span = np.maximum(hi - lo, 1e-6)
    return (lo - margin * span).astype(np.float32), (hi + margin * span).astype(np.float32)


def train(r: plugin.Renderer, n: int, epochs: int, batch: int, lr: float, seed: int, log=print) -> dict:
    """Adam on the collected samples; returns the fitted range and loss scale."""
    x, t = r.mlp_read_samples(n)
    scale = (1.0 / np.maximum(t.std(axis=0), 1e-4)).astype(np.float32)
    lo, hi = feature_range(x)
    loss0, _, _ = r.mlp_gradients(x[:8192], t[:8192], scale)
    t0 = time.perf_counter()
    last = None
    for ep in range(epochs):
        step_lr = lr * (0.5 * (1.0 + np.cos(np.pi * ep / max(1, epochs))))  # cosine decay
        last = r.mlp_train(1, batch, float(step_lr), scale, seed + ep)
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            log(f"[mlp] epoch {ep}: loss {last:.5f} (lr {step_lr:.2e}, {time.perf_counter() - t0:.1f}s)")
    r.mlp_set_range(lo, hi)
    held = slice(0, min(n, 65536))
    y = r.mlp_forward(x[held])[:, :N.MLP_NOUT]
    rel = np.sqrt(((y - t[held]) ** 2).mean(axis=0)) / np.maximum(np.abs(t[held]).mean(axis=0), 1e-9)
    log("[mlp] relative RMS error per output: " + " ".join(f"{v:.3f}" for v in rel))
    return {"lo": lo, "hi": hi, "scale": scale, "loss0": float(loss0), "loss": float(last), "samples": n,
            "rel_rms": rel}


def save(path: str, r: plugin.Renderer, info: dict):
    w, b = r.mlp_get_weights()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=w, biases=b, lo=info["lo"], hi=info["hi"], scale=info["scale"],
             samples=np.int64(info["samples"]), loss=np.float32(info["loss"]))


def without_term(r: plugin.Renderer):
    """A network whose outputs decode to zero dual-scattering terms: shading
    without multiple scattering (the ablation the trained network is compared
    with)."""
    w = np.zeros(N.MLP_NW, np.float32)
    b = np.zeros(N.MLP_NBIAS, np.float32)
    out = np.arange(N.MLP_NOUT)
    log_out = (out < 10) | (out >= 13)  # outputs predicted as log(value + LOG_EPS)
    b[192 + out[log_out]] = np.log(LOG_EPS)
    r.mlp_set_weights(w, b)
    r.mlp_set_range(None, None)


def install(r: plugin.Renderer, path: str = DEFAULT_WEIGHTS) -> dict:
    """Loads trained weights and their feature range into the renderer."""
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if d["scale"].shape != (N.MLP_NOUT,):
        raise ValueError(f"{path}: weights for {d['scale'].shape[0]} outputs, the network has {N.MLP_NOUT}; retrain")
    r.mlp_set_weights(d["weights"], d["biases"])
    r.mlp_set_range(d["lo"], d["hi"])
    return d


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--workloads", default="c1,c2")
    ap.add_argument("--directions", type=int, default=64)
    ap.add_argument("--samples", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)
    r = plugin.Renderer(0)
    r.mlp_init(args.seed)
    dirs = key_directions(args.directions, args.seed)
    n = collect(r, args.workloads.split(","), dirs, args.samples, args.seed)
    info = train(r, n, args.epochs, args.batch, args.lr, args.seed)
    save(args.out, r, info)
    print(f"[mlp] saved {args.out}: {n} samples, loss {info['loss0']:.4f} -> {info['loss']:.5f}", flush=True)
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
