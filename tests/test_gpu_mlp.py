"""The learned shading term on the GPU: forward and gradients against the
float64 oracle (oracle/mlp.py), training, and MLP shading against the plain
method with the per-tile fallback."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as M
from bevy_hair_amd import _native as N

pytestmark = pytest.mark.gpu


def unpack(w, b):
    ws, off = [], 0
    for i in range(4):
        o, k = M.DIMS[i + 1], M.DIMS[i]
        ws.append(w[off:off + o * k].reshape(o, k))
        off += o * k
    bs = [b[0:64], b[64:128], b[128:192], b[192:208]]
    return ws, bs


def random_x(rng, n):
    x = np.zeros((n, 32), np.float32)
    x[:, :M.N_IN] = rng.uniform(-1.0, 1.0, (n, M.N_IN))
    return x


def test_forward_matches_fp64_oracle(renderer):
    print("\n[gpu mlp] forward", flush=True)
    renderer.mlp_init(7)
    w, b = renderer.mlp_get_weights()
    rng = np.random.default_rng(0)
    b = rng.normal(0, 0.1, b.shape).astype(np.float32)
    b[192 + M.N_OUT:] = 0.0  # padded outputs keep zero biases
    renderer.mlp_set_weights(w, b)
    x = random_x(rng, 1000)
    y = renderer.mlp_forward(x)
    ws, bs = unpack(w, b)
    ref, _ = M.forward(ws, bs, x, bf16_operands=True)
    if M.N_OUT < 16:
        assert np.abs(y[:, M.N_OUT:]).max() == 0.0  # padded outputs stay zero
    err = np.abs(y[:, :M.N_OUT] - ref[:, :M.N_OUT])
    scale = np.abs(ref[:, :M.N_OUT]).max()
    # f32 accumulation and the rounding of hidden activations to bf16 from f32
    # instead of f64 move an output by a few bf16 ulps at most
    assert err.max() <= 1e-2 * scale, (err.max(), scale)
    assert np.median(err) <= 1e-3 * scale


def test_gradients_match_fp64_oracle(renderer):
    print("\n[gpu mlp] gradients", flush=True)
    renderer.mlp_init(3)
    w, b = renderer.mlp_get_weights()
    rng = np.random.default_rng(1)
    x = random_x(rng, 512)
    t = rng.normal(0, 0.5, (512, M.N_OUT)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, M.N_OUT).astype(np.float32)
    loss, dw, db = renderer.mlp_gradients(x, t, scale)
    ws, bs = unpack(w, b)
    ref_loss, rdw, rdb = M.loss_and_grads(ws, bs, x, t, scale, bf16_operands=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    got = unpack(dw, db)
    for L in range(4):
        g, r = got[0][L].astype(np.float64), rdw[L]
        cos = (g * r).sum() / (np.linalg.norm(g) * np.linalg.norm(r))
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"layer {L}: dW cos {cos:.6f} rel {rel:.2e}")
        assert cos > 0.9995 and rel < 3e-2
        gb, rb = got[1][L].astype(np.float64), rdb[L]
        assert np.linalg.norm(gb - rb) <= 3e-2 * np.linalg.norm(rb) + 1e-12
    # padded weights get no gradient
    if M.N_IN < 32:
        assert np.abs(got[0][0][:, M.N_IN:]).max() == 0.0
    if M.N_OUT < 16:
        assert np.abs(got[0][3][M.N_OUT:, :]).max() == 0.0


def test_training_fits_a_smooth_teacher(renderer):
    """Adam through the GPU's own forward/backward kernels on a fixed target function."""
    print("\n[gpu mlp] training", flush=True)
    rng = np.random.default_rng(5)
    n = 1 << 15
    x = random_x(rng, n)
    proj = rng.normal(0, 1.0, (M.N_IN, M.N_OUT))
    t = np.tanh(x[:, :M.N_IN] @ proj * 0.3).astype(np.float32)
    renderer.mlp_init(11)
    renderer.mlp_reserve_samples(n)
    scale = np.ones(M.N_OUT, np.float32)
    loss0, _, _ = renderer.mlp_gradients(x[:4096], t[:4096], scale)
    renderer.mlp_set_samples(x, t)
    last = renderer.mlp_train(epochs=30, batch=2048, lr=3e-3, scale=scale, seed=2)
    loss1, _, _ = renderer.mlp_gradients(x[:4096], t[:4096], scale)
    print(f"loss {loss0:.4f} -> {loss1:.5f} (last epoch {last:.5f})")
    assert loss1 < 0.1 * loss0



# ---- network shading --------------------------------------------------------------


def _frame(renderer, name):
    """Rasterizes the named workload (G-buffer + DOM) and returns (scene, groom)."""
    from bevy_hair_amd.workload import make_scene
    sc = make_scene(name)
    w = sc.w
    renderer.targets(w.width, w.height, w.dom_size)
    renderer.set_scene_depth(None, 0.0)
    renderer.set_lights(sc.points, sc.dirs)
    g = renderer.upload(sc.mesh)
    opts = renderer.frame_opts(control_points=w.control_points, strand_fraction=w.strand_fraction)
    renderer.hair_pass(g, sc.params, opts)
    return sc, g


def _rel_err(got, ref, floor=1e-3):
    """Mean over covered pixels of |got - ref|_1 / max(|ref|_1, floor) (RGB)."""
    cov = ref[:, 3] > 0
    d = np.abs(got[cov, :3] - ref[cov, :3]).sum(1)
    return float((d / np.maximum(np.abs(ref[cov, :3]).sum(1), floor)).mean())


def test_out_of_range_tiles_are_the_plain_method(renderer):
    """With an empty trained range every tile falls back: the network path's
    image is the plain kernel's, bit for bit, and every tile is counted."""
    print("\n[gpu mlp] fallback = plain (c1)", flush=True)
    from bevy_hair_amd import mlp_train
    mlp_train.install(renderer)
    sc, g = _frame(renderer, "c1")
    try:
        renderer.shade(sc.params[0])
        plain_img, plain_depth = (a.copy() for a in renderer.shaded())
        lo = np.full(32, 1e30, np.float32)
        hi = np.full(32, -1e30, np.float32)
        renderer.mlp_set_range(lo, hi)
        n = renderer.shade_mlp(sc.params[0], 1.0, 0.0)
        img, depth = renderer.shaded()
        covered = int((plain_img[:, 3] > 0).sum())
        assert n == (covered + 63) // 64
        assert np.array_equal(img, plain_img) and np.array_equal(depth, plain_depth)
    finally:
        mlp_train.install(renderer)
        g.free()


@pytest.mark.parametrize("name", ["c4", "c2"])
def test_trained_network_error_and_gain(renderer, name):
    """The committed network (trained on c1 + c2 renders under 64 random key
    directions): mean relative RGB error against the plain method at most 2%
    with the fallback, and far below the error of shading without the learned
    terms (the ablation)."""
    print(f"\n[gpu mlp] trained network ({name})", flush=True)
    from bevy_hair_amd import mlp_train
    mlp_train.install(renderer)
    sc, g = _frame(renderer, name)
    try:
        renderer.shade(sc.params[0])
        ref = renderer.shaded()[0].copy()
        fallback = renderer.shade_mlp(sc.params[0], 1.0, 0.05)
        with_fb = _rel_err(renderer.shaded()[0], ref)
        renderer.mlp_set_range(None, None)
        renderer.shade_mlp(sc.params[0], 1.0, 0.0)
        raw = _rel_err(renderer.shaded()[0], ref)
        mlp_train.without_term(renderer)
        renderer.shade_mlp(sc.params[0], 1.0, 0.0)
        without = _rel_err(renderer.shaded()[0], ref)
        print(f"{name}: mean rel err with fallback {with_fb:.4f} ({fallback} plain tiles), network only {raw:.4f}, "
              f"without the learned terms {without:.4f}")
        assert with_fb <= 0.02
        assert raw <= 0.5 * without
    finally:
        mlp_train.install(renderer)
        g.free()
