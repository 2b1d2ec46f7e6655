"""Frames/s of the deferred strand-rasterization path on one MI355X.

A step is one full hair_pass of one groom over resident inputs: clear, camera
LOD/scan/raster, DOM clear/LOD/scan/raster x2, shade, filter. The default
workload is the 1080p, 100k-strand x 64-segment wavy groom with a 2048^2 x 16
deep opacity map and 1 directional + 32 point lights. `value` is the plain
method's frames/s; the "mlp" entry times the same frames with the learned
dual-scattering term and gives its error against the plain method, and the
"batch" entry the views/s of a 24-view, three-groom batch replayed from one
captured HIP graph.

Prints one JSON line (rank 0). Per-kernel HIP events are recorded on the
library's own stream across the timed frames; the dominant kernel's mean
duration gives the roofline entry. The CPU baseline (the build's C++
restatement, multithreaded) is timed on a bounded sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "1080p frames/s (100k×64 groom): plain vs MLP shading at ≤2% err; HBM GB/s"
HBM_PEAK_GBS = 8000.0


# Timer slots that hold one kernel launch each (the roofline's candidates).
SINGLE_KERNELS = ("clear", "cam_raster", "clear_dom", "dom_depth", "dom_opacity", "shade", "filter", "dom_tiles",
                  "cam_tiles")


def algorithmic_bytes(kernel: str, w, mesh, counts: dict, hit_pixels: int) -> float:
    """Algorithmic HBM bytes of one launch by the path's own work (per-unit
    figures x units): clears 20 B per pixel and 68 B per DOM texel; 8 B per
    u64 and 4 B per u32 atomic, each counted once; 4 B of scene depth per
    camera write; the groom's compulsory bytes once per pass; shade
    (8 + 4 + 20 + 272) B per covered pixel; filter 56 B per pixel."""
    P = w.width * w.height
    S = w.dom_size
    B, N = mesh.bundle_count, mesh.strand_count
    groom = B * (16 + 64 * 32 + 1024 * 16) + 16 * N
    cam_raster = 8.0 * counts["cam"][0] + 12.0 * counts["cam"][1] + 4.0 * counts["cam"][0]
    dom_atomics = 4.0 * counts["dom1"][2] + 4.0 * counts["dom2"][2]
    if kernel == "clear":
        return 20.0 * P
    if kernel == "clear_dom":
        return 68.0 * S * S
    if kernel == "cam_raster":
        return cam_raster + groom
    if kernel == "dom_depth":
        return 4.0 * counts["dom1"][2] + groom
    if kernel == "dom_opacity":
        return 4.0 * counts["dom2"][2]
    if kernel == "cam_tiles":  # the camera raster's atomics and the G-buffer clear it performs
        return 20.0 * P + cam_raster
    if kernel == "dom_tiles":  # both DOM passes' atomics and the DOM clear it performs
        return 68.0 * S * S + dom_atomics
    if kernel == "shade":
        return (8.0 + 4.0 + 20.0 + 272.0) * hit_pixels
    if kernel == "filter":
        return 56.0 * P
    return 0.0


def implementation_bytes(kernel: str, w, mesh, counts: dict, hit_pixels: int) -> float:
    """HBM bytes the binned implementation moves by design (its own control
    points, tile boxes and records on top of the algorithmic bytes)."""
    P = w.width * w.height
    S = w.dom_size
    B, N = mesh.bundle_count, mesh.strand_count
    groom = B * (16 + 64 * 32 + 1024 * 16) + 16 * N
    if kernel == "cam_bin":
        # groom in; per control point world position + screen position (20 B)
        # and world tangent + ao (16 B) out; 16 B per strand; tile box out and
        # back per segment; one 4 B record per (segment, tile)
        return (groom + 36.0 * counts["cam_cps"] + 16.0 * N + 8.0 * counts["cam_segments"]
                + 4.0 * counts["cam_records"])
    if kernel == "cam_tiles":
        # the G-buffer written once (centre, conservative, beta); per record its
        # id, both control points (36 B each) and the strand's uv / beta
        return 20.0 * P + 92.0 * counts["cam_records"]
    if kernel == "dom_bin":
        # groom in, control points out (20 B), tile box out and back (4 + 4 B
        # per segment), one 4 B record per (segment, tile)
        return groom + 20.0 * counts["dom_cps"] + 8.0 * counts["dom_segments"] + 4.0 * counts["dom_records"]
    if kernel == "dom_tiles":
        # the map written once (front + 16 counts per texel); per record its
        # id and its two 20-byte control points
        return 68.0 * S * S + 44.0 * counts["dom_records"]
    if kernel == "shade":
        return 8.0 * P + 20.0 * P + hit_pixels * (4.0 + 272.0)
    return algorithmic_bytes(kernel, w, mesh, counts, hit_pixels)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """The host cores this process may use (its affinity set), at most 16:
    the CPU share of one GPU on the measuring machine."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(sc, w, threads: int, fraction: int) -> dict:
    """The oracle on 1/fraction of the frame's work, scaled back to a full frame."""
    import oracle

    ref = oracle.Frame(sc.mesh, w.width, w.height, w.dom_size)
    cap = max(1, sc.mesh.strand_count)
    cam = sc.params[0]
    pa, n_p, da, n_d = sc.light_arrays()
    t = {}
    t0 = time.perf_counter()
    ref.clear()
    lod = ref.lod(cam, w.control_points, w.strand_fraction)
    refs, _ = ref.scan(cam, lod, cap)
    t["setup"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref.raster(cam, refs[::fraction].copy(), threads=threads)
    t["cam_raster"] = (time.perf_counter() - t0) * fraction
    ref.clear_dom()
    dlod = ref.lod(sc.params[1])
    drefs, _ = ref.scan(sc.params[1], dlod, cap)
    sub = drefs[::fraction].copy()
    t0 = time.perf_counter()
    ref.raster(sc.params[1], sub, threads=threads)
    ref.raster(sc.params[2], sub, threads=threads)
    t["dom"] = (time.perf_counter() - t0) * fraction
    # shade and filter a band of rows through the middle of the image
    r0 = w.height // 2 - w.height // (2 * fraction)
    rows = (r0, r0 + w.height // fraction)
    t0 = time.perf_counter()
    ref.shade(cam, pa, n_p, da, n_d, threads=threads, rows=rows)
    t["shade"] = (time.perf_counter() - t0) * fraction
    t0 = time.perf_counter()
    ref.filter(cam, threads=threads, rows=rows)
    t["filter"] = (time.perf_counter() - t0) * fraction
    frame_s = sum(t.values())
    return {"value": round(1.0 / frame_s, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ restatement, not lavapipe) on every {fraction}th strand of the camera and "
                      f"DOM passes and {rows[1] - rows[0]} of {w.height} rows of shade+filter, scaled x{fraction}",
            "seconds_per_frame": {k: round(v, 3) for k, v in t.items()}}


def cpu_baseline_report(sc, w, threads: int, fraction: int) -> dict:
    """All-core run on 1/fraction of the frame plus a one-thread run on a
    smaller sample (1/(4 fraction)), both scaled to a full frame."""
    out = cpu_baseline(sc, w, threads, fraction)
    one = cpu_baseline(sc, w, 1, 4 * fraction)
    out["cpu_model"] = cpu_model()
    out["one_thread"] = {"value": one["value"], "sample": one["sample"], "seconds_per_frame": one["seconds_per_frame"]}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c2_traffic_pmc.txt")


def measured_traffic(kernel: str, tuning: dict, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by
    what the counter reads for a 1 GiB copy kernel); None when the summary
    does not cover this kernel on this workload."""
    dom_name = ("k_dom_tiles_busy" if tuning.get("dom_tile") == 16 and tuning.get("dom_busy_blocks_per_cu")
                else f"k_dom_tiles<{tuning.get('dom_tile')}>")
    names = {"cam_tiles": f"k_cam_tiles<{tuning.get('cam_tile')}>", "dom_tiles": dom_name,
             "shade": "k_shade<0u>", "filter": "k_filter"}
    if kernel not in names or not os.path.exists(TRAFFIC_FILE):
        return None
    rows, wl = {}, None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) >= 2 and f[0] == "workload":
            wl = f[1]
        elif len(f) == 3 and not line.startswith("#"):
            try:
                rows[f[0]] = (float(f[1]), float(f[2]))
            except ValueError:
                pass
    if wl != workload or "copy_1GiB" not in rows or names[kernel] not in rows:
        return None
    cf, cw = rows["copy_1GiB"]
    sf, sw = 1048576.0 / cf, 1048576.0 / cw  # the copy reads and writes 1 GiB
    fk, wk = rows[names[kernel]]
    return round((fk * sf + wk * sw) * 1024.0)


def rel_err(got: np.ndarray, ref: np.ndarray, floor: float = 1e-3) -> float:
    """Mean over covered pixels of |got - ref|_1 / max(|ref|_1, floor) over RGB."""
    cov = ref[:, 3] > 0
    d = np.abs(got[cov, :3] - ref[cov, :3]).sum(1)
    return float((d / np.maximum(np.abs(ref[cov, :3]).sum(1), floor)).mean())


def timed_frames(r, g, params, opts, steps: int):
    """K frames between one event pair with nothing else on the stream (the
    throughput, the DOM pass beside the camera pass as the launch choices
    say), then K more with an event pair around every kernel group and the
    two passes one after the other (the per-kernel durations, in ms per
    frame, each kernel with the GPU to itself): the per-kernel events
    themselves cost a few microseconds per frame, so they stay out of the
    first measurement. Returns (ms per frame, ms per frame with the kernel
    events, kernel times)."""
    r.sync()
    ms = r.time_frames(g, params, opts, steps) / steps
    keep = r.get_tuning()["frame_serial"]
    r.set_tuning(frame_serial=1)
    try:
        r.timing(steps)
        r.sync()
        ms_events = r.time_frames(g, params, opts, steps) / steps
        kt = r.kernel_times()
        r.timing(0)
    finally:
        r.set_tuning(frame_serial=keep)
    return ms, ms_events, kt


def one_stream_ms(r, g, params, opts, steps: int) -> float:
    """ms per frame with the DOM pass after the camera pass on one stream."""
    keep = r.get_tuning()["frame_serial"]
    r.set_tuning(frame_serial=1)
    try:
        r.sync()
        return r.time_frames(g, params, opts, steps) / steps
    finally:
        r.set_tuning(frame_serial=keep)


def mlp_leg_run(r, g, sc, w, args) -> dict:
    """The same frames with the learned dual-scattering term (committed
    weights): frames/s, shade time against the plain shade, and the mean
    relative RGB error of the shaded image against the plain method's."""
    from bevy_hair_amd import mlp_train

    info = mlp_train.install(r)
    mopts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction, mlp=True,
                         mlp_tolerance=args.mlp_tolerance)
    popts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    for _ in range(args.warmup):
        r.hair_pass(g, sc.params, mopts)
    ms, _, kt = timed_frames(r, g, sc.params, mopts, args.steps)
    r.hair_pass(g, sc.params, popts)
    r.shade(sc.params[0])
    plain = r.shaded()[0].copy()
    fallback = r.shade_mlp(sc.params[0], 1.0, args.mlp_tolerance)
    err = rel_err(r.shaded()[0], plain)
    return {"frames_per_s": round(1000.0 / ms, 2), "ms_per_step": round(ms, 4), "shade_ms": round(kt["shade"], 5),
            "mean_rel_rgb_err": round(err, 5), "fallback_tiles": fallback, "tolerance": args.mlp_tolerance,
            "dtype": "bf16 MFMA, f32 accumulate", "training_samples": int(info.get("samples", 0))}


def mlp_c4_leg_run(r, args) -> dict:
    """The learned-shading configuration's evaluation view: the config-1 groom
    (4096 straight strands x 16 segments) at 1080p with one shadow-casting
    directional light. Frames/s and shade time with the plain method and with
    the learned term, and the learned image's mean relative RGB error."""
    from bevy_hair_amd import mlp_train
    from bevy_hair_amd.workload import make_scene

    sc = make_scene("c4")
    w = sc.w
    mlp_train.install(r)
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    out = {"workload": w.name, "lights": len(sc.points) + len(sc.dirs)}
    try:
        steps = max(10, args.steps)
        for name, mlp in (("plain", False), ("mlp", True)):
            opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction, mlp=mlp,
                                mlp_tolerance=args.mlp_tolerance)
            for _ in range(3):
                r.hair_pass(g, sc.params, opts)
            ms, _, kt = timed_frames(r, g, sc.params, opts, steps)
            out[name] = {"frames_per_s": round(1000.0 / ms, 2), "shade_ms": round(kt["shade"], 5)}
        r.hair_pass(g, sc.params, r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction))
        r.shade(sc.params[0])
        plain = r.shaded()[0].copy()
        out["fallback_tiles"] = r.shade_mlp(sc.params[0], 1.0, args.mlp_tolerance)
        out["mean_rel_rgb_err"] = round(rel_err(r.shaded()[0], plain), 5)
    finally:
        g.free()
    return out


def batch_leg_run(r, args) -> dict:
    """24 views over three grooms (20k x 16, 100k x 64, 150k x 126) at
    1280x720 / 1920x1080 / 2560x1440, learned shading with the trained-range
    fallback, replayed from one captured HIP graph: views/s with the views
    dealt to the tuned number of lanes, and with one lane beside it."""
    from bevy_hair_amd import mlp_train
    from bevy_hair_amd import workload as W

    bs = W.make_batch_scene()
    mlp_train.install(r)
    r.set_lights(bs.points, bs.dirs)
    gpu = [r.upload(m) for m in bs.meshes]
    views = [(gpu[gi], params, wd, ht, W.BATCH_DOM,
              r.frame_opts(control_points=ncp, strand_fraction=1.0, mlp=True, mlp_tolerance=args.mlp_tolerance))
             for gi, wd, ht, _, params, ncp in bs.views]
    reps = max(3, min(20, args.steps // 5))
    lanes = r.get_tuning()["batch_lanes"]
    per_lanes = {}
    try:
        for n in sorted({lanes, 1}, reverse=True):
            r.set_tuning(batch_lanes=n)
            b = r.batch(views)
            try:
                b.run(2)
                per_lanes[n] = b.time(reps) / reps
            finally:
                b.free()
    finally:
        r.set_tuning(batch_lanes=lanes)
        for gg in gpu:
            gg.free()
    ms = per_lanes[lanes]
    return {"views": len(views), "views_per_s": round(len(views) * 1000.0 / ms, 1), "ms_per_batch": round(ms, 3),
            "lanes": lanes, "views_per_s_one_lane": round(len(views) * 1000.0 / per_lanes[1], 1),
            "replays": reps, "grooms": [f"{k} {n}x{ncp - 1}" for k, n, _, ncp in W.BATCH_GROOMS],
            "sizes": [f"{a}x{b_}" for a, b_ in W.BATCH_SIZES], "dom_size": W.BATCH_DOM, "shading": "mlp"}


def heavy_leg_run(r, args) -> dict:
    """The 2160p heavy workload (60k curly strands x 126 segments, DOM 2048^2,
    1 + 128 lights) at camera LOD strand fractions 1, 1/2 and 1/4, plain
    shading, with the launch choices in force: frames/s per fraction."""
    from bevy_hair_amd.workload import make_scene

    sc = make_scene("c3")
    w = sc.w
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    out = {"workload": w.name, "point_lights": len(sc.points), "fractions": {}}
    try:
        steps = max(5, args.steps // 5)
        for frac in (1.0, 0.5, 0.25):
            opts = r.frame_opts(control_points=w.control_points, strand_fraction=frac)
            for _ in range(2):
                r.hair_pass(g, sc.params, opts)
            ms, _, kt = timed_frames(r, g, sc.params, opts, steps)
            top = sorted(((k, v) for k, v in kt.items() if v), key=lambda kv: -kv[1])[:4]
            out["fractions"][str(frac)] = {"frames_per_s": round(1000.0 / ms, 2), "ms_per_step": round(ms, 4),
                                           "top_kernels_ms": {k: round(v, 4) for k, v in top}}
    finally:
        g.free()
    return out


def lights_leg_run(r, g, sc, w, args) -> dict:
    """The same groom and view under 1 key + 1024 short-range point and spot
    lights, three rect lights and a studio probe: frames/s and shade time
    with the lights culled per wave (the automatic choice here) and without."""
    from bevy_hair_amd import scene

    centre = 0.5 * (sc.mesh.bounds_min + sc.mesh.bounds_max).astype(np.float64)
    radius = 0.5 * float(np.linalg.norm(sc.mesh.bounds_max - sc.mesh.bounds_min))
    pts = scene.scattered_lights(np.random.default_rng(5), 1024, centre, radius, 0.15, 0.6)
    keep = r.get_tuning()
    opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    out = {"point_and_spot_lights": len(pts), "rect_lights": 3, "probe": "64^2 x 7 specular, 16^2 diffuse"}
    try:
        r.set_lights(pts, sc.dirs)
        r.set_rect_lights(scene.studio_rects(centre, radius))
        r.set_environment(scene.studio_environment(64, 7, 16, intensity=0.8))
        steps = max(5, args.steps // 5)
        for name, cull in (("culled_per_wave", 0), ("every_light_per_pixel", 1)):
            r.set_tuning(light_cull=cull)
            for _ in range(2):
                r.hair_pass(g, sc.params, opts)
            ms, _, kt = timed_frames(r, g, sc.params, opts, steps)
            out[name] = {"frames_per_s": round(1000.0 / ms, 2), "shade_ms": round(kt["shade"], 4)}
    finally:
        r.set_tuning(**keep)
        r.set_rect_lights([])
        r.set_environment(None)
        r.set_lights(sc.points, sc.dirs)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-fraction", type=int, default=2)
    ap.add_argument("--no-probes", action="store_true", help="skip the copy/atomic microbenchmarks")
    ap.add_argument("--no-tune", action="store_true", help="keep the default raster launch choices")
    ap.add_argument("--tuning", default=None,
                    help="launch choices as JSON (hm_tuning fields) instead of autotuning, e.g. for profiling")
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned-shading legs")
    ap.add_argument("--no-mlp-c4", action="store_true",
                    help="skip the learned-shading leg on the config-1 groom at 1080p")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-view leg")
    ap.add_argument("--no-heavy", action="store_true", help="skip the 2160p heavy-workload leg")
    ap.add_argument("--no-lights", action="store_true", help="skip the many-lights / probe leg")
    ap.add_argument("--mlp-tolerance", type=float, default=0.0,
                    help="per-tile probe bound of the learned shading (0: trained-range fallback only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this path runs on a single MI355X (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    from bevy_hair_amd import _native as N
    from bevy_hair_amd import plugin
    from bevy_hair_amd.workload import make_scene

    t_setup = time.perf_counter()
    sc = make_scene(args.workload)
    w = sc.w
    print(f"[bench] workload {w.name}: {sc.mesh.strand_count} strands in {sc.mesh.bundle_count} bundles, "
          f"setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    r = plugin.Renderer(0)
    r.targets(w.width, w.height, w.dom_size)
    r.set_scene_depth(None, 0.0)
    r.set_lights(sc.points, sc.dirs)
    g = r.upload(sc.mesh)
    opts = r.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    for _ in range(args.warmup):
        r.hair_pass(g, sc.params, opts)
    r.sync()
    # Launch choices of the raster passes picked on this groom and view (untimed).
    if args.tuning:
        r.set_tuning(**json.loads(args.tuning))
        tuning = r.get_tuning()
    else:
        tuning = r.autotune(g, sc.params, opts, reps=3) if not args.no_tune else r.get_tuning()
    r.hair_pass(g, sc.params, opts)
    r.sync()
    # Work counts of each raster pass (untimed counting launches).
    r.clear()
    r.lod(g, sc.params[0], N.PASS_CAMERA, w.control_points, w.strand_fraction)
    r.scan(g, sc.params[0], N.PASS_CAMERA)
    counts = {"cam": r.raster_count(g, sc.params[0], N.PASS_CAMERA)}
    cam_refs, _ = r.refs(g, N.PASS_CAMERA)
    cncp = np.clip(cam_refs[:, 3].astype(np.int64), 2, 127)
    counts["cam_cps"] = int(cncp.sum())
    counts["cam_segments"] = int((cncp - 1).sum())
    r.gbuffer_pass(g, sc.params[0])
    cam_status = r.bin_status(0) if tuning.get("cam_tile") else None
    counts["cam_records"] = cam_status["records"] if cam_status else 0
    r.hair_pass(g, sc.params, opts)
    r.lod(g, sc.params[1], N.PASS_DOM)
    r.scan(g, sc.params[1], N.PASS_DOM)
    counts["dom1"] = r.raster_count(g, sc.params[1], N.PASS_DOM)
    counts["dom2"] = r.raster_count(g, sc.params[2], N.PASS_DOM)
    dom_refs, _ = r.refs(g, N.PASS_DOM)
    ncp = np.clip(dom_refs[:, 3].astype(np.int64), 2, 127)
    counts["dom_cps"] = int(ncp.sum())
    counts["dom_segments"] = int((ncp - 1).sum())
    r.dom(g, sc.params[1], sc.params[2])
    bin_status = r.bin_status(1) if tuning.get("dom_tile") else None
    counts["dom_records"] = bin_status["records"] if bin_status else 0
    center, _, _ = r.gbuffer()
    hit = int((center != np.uint64(0xFFFFFFFFFFFFFFFF)).sum())
    r.sync()

    # Timed region: K frames between two events; then K frames with the
    # per-kernel events that give the kernel breakdown and the roofline.
    ms, ms_events, kt = timed_frames(r, g, sc.params, opts, args.steps)
    ms_serial = one_stream_ms(r, g, sc.params, opts, args.steps)
    fps = 1000.0 / ms

    mlp_leg = None if args.no_mlp else mlp_leg_run(r, g, sc, w, args)
    # the dominant single kernel (the bin groups time four kernels each)
    dom_k = max((k for k in kt if k in SINGLE_KERNELS), key=kt.get)
    alg = algorithmic_bytes(dom_k, w, sc.mesh, counts, hit)
    impl = implementation_bytes(dom_k, w, sc.mesh, counts, hit)
    achieved = alg / (kt[dom_k] * 1e-3) / 1e9
    atomics_cam = float(counts["cam"][0] + 2 * counts["cam"][1])
    atomics_dom = float(counts["dom1"][2] + counts["dom2"][2])
    per_kernel = {k: {"ms": round(kt[k], 5), "algorithmic_GBs": round(algorithmic_bytes(k, w, sc.mesh, counts, hit)
                                                                       / (kt[k] * 1e-3) / 1e9, 1)}
                  for k in ("cam_tiles", "dom_tiles", "shade", "filter") if kt.get(k)}
    result = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (u64/u32 integer keys and counts)",
        "data": "synthetic",
        "config": {"workload": w.name, "strands": sc.mesh.strand_count, "bundles": sc.mesh.bundle_count,
                   "segments_per_strand": w.control_points - 1, "width": w.width, "height": w.height,
                   "dom_size": w.dom_size, "dom_layers": 16, "point_lights": len(sc.points),
                   "directional_lights": len(sc.dirs), "shading": "plain", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_k, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "algorithmic_bytes": alg, "implementation_bytes": impl, "mean_ms": round(kt[dom_k], 5),
                     "traffic": measured_traffic(dom_k, tuning, w.name)},
        "kernel_rooflines": per_kernel,
        "kernels_ms": {k: round(v, 5) for k, v in kt.items()},
        "ms_per_step_with_kernel_events": round(ms_events, 4),
        "ms_per_step_one_stream": round(ms_serial, 4),
        "tuning": tuning,
        "work": {"camera_writes": int(counts["cam"][0]), "camera_center_hits": int(counts["cam"][1]),
                 "camera_atomics": int(atomics_cam), "dom_depth_writes": int(counts["dom1"][2]),
                 "dom_opacity_increments": int(counts["dom2"][2]), "segments": int(counts["cam"][3]),
                 "hit_pixels": hit, "cam_bin": cam_status, "dom_bin": bin_status},
        "atomic_rates": {
            "cam_raster_atomics_per_s": round(atomics_cam / (kt["cam_raster"] * 1e-3), 1)
            if kt["cam_raster"] and not cam_status else None,
            "dom_opacity_atomics_per_s": round(float(counts["dom2"][2]) / (kt["dom_opacity"] * 1e-3), 1)
            if kt["dom_opacity"] and not bin_status else None,
            # the binned path's atomics run in LDS: the camera tiles' u64 / u32
            # minima and the DOM tiles' u32 minima and adds, per second of kernel
            "cam_tiles_lds_atomics_per_s": round(atomics_cam / (kt["cam_tiles"] * 1e-3), 1)
            if kt.get("cam_tiles") else None,
            "dom_tiles_lds_atomics_per_s": round(atomics_dom / (kt["dom_tiles"] * 1e-3), 1)
            if kt.get("dom_tiles") else None},
    }
    result["mlp"] = mlp_leg
    if not args.no_probes:
        # Measured denominators: STREAM copy and random-address atomics over the
        # footprints the raster passes hit (G-buffer words, DOM fronts, DOM counts).
        P, S = w.width * w.height, w.dom_size
        peaks = {"copy_GBs": round(r.bench_copy(1 << 30, 5), 1),
                 "u64_min_per_s": r.bench_atomics(8 * P, 0, 1 << 26),
                 "u32_min_per_s": r.bench_atomics(4 * S * S, 1, 1 << 26),
                 "u32_add_per_s": r.bench_atomics(64 * S * S, 2, 1 << 26),
                 # LDS atomics over tile-sized arrays: 32x32 u64 words, 16x16x16 u32 counts
                 "lds_u64_min_per_s": r.bench_lds_atomics(8192, 0, 1 << 30),
                 "lds_u32_min_per_s": r.bench_lds_atomics(16384, 1, 1 << 30),
                 "lds_u32_add_per_s": r.bench_lds_atomics(16384, 2, 1 << 30)}
        result["measured_peaks"] = peaks
        rates = result["atomic_rates"]
        if rates["cam_tiles_lds_atomics_per_s"]:
            rates["cam_tiles_frac_of_lds_u64_min_peak"] = round(
                rates["cam_tiles_lds_atomics_per_s"] / peaks["lds_u64_min_per_s"], 4)
        if rates["dom_tiles_lds_atomics_per_s"]:
            rates["dom_tiles_frac_of_lds_u32_add_peak"] = round(
                rates["dom_tiles_lds_atomics_per_s"] / peaks["lds_u32_add_per_s"], 4)
        if rates["cam_raster_atomics_per_s"]:
            rates["cam_raster_frac_of_u64_min_peak"] = round(rates["cam_raster_atomics_per_s"] / peaks["u64_min_per_s"], 4)
        if rates["dom_opacity_atomics_per_s"]:
            rates["dom_opacity_frac_of_u32_add_peak"] = round(rates["dom_opacity_atomics_per_s"] / peaks["u32_add_per_s"], 4)
    result["many_lights"] = None if args.no_lights else lights_leg_run(r, g, sc, w, args)
    g.free()
    if result["mlp"] is not None and not args.no_mlp_c4:
        result["mlp"]["c4"] = mlp_c4_leg_run(r, args)
    result["batch"] = None if args.no_batch else batch_leg_run(r, args)
    result["heavy"] = None if args.no_heavy else heavy_leg_run(r, args)
    r.close()
    if not args.no_cpu:
        threads = host_threads()
        print(f"[bench] cpu baseline on {threads} threads ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline_report(sc, w, threads, args.cpu_fraction)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
