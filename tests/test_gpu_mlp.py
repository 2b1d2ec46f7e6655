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
    x[:, :29] = rng.uniform(-1.0, 1.0, (n, 29))
    return x


def test_forward_matches_fp64_oracle(renderer):
    print("\n[gpu mlp] forward", flush=True)
    renderer.mlp_init(7)
    w, b = renderer.mlp_get_weights()
    rng = np.random.default_rng(0)
    b = rng.normal(0, 0.1, b.shape).astype(np.float32)
    b[204:] = 0.0
    renderer.mlp_set_weights(w, b)
    x = random_x(rng, 1000)
    y = renderer.mlp_forward(x)
    ws, bs = unpack(w, b)
    ref, _ = M.forward(ws, bs, x, bf16_operands=True)
    assert np.abs(y[:, 12:]).max() == 0.0  # padded outputs stay zero
    err = np.abs(y[:, :12] - ref[:, :12])
    scale = np.abs(ref[:, :12]).max()
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
    t = rng.normal(0, 0.5, (512, 12)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, 12).astype(np.float32)
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
    assert np.abs(got[0][0][:, 29:]).max() == 0.0
    assert np.abs(got[0][3][12:, :]).max() == 0.0


def test_training_fits_a_smooth_teacher(renderer):
    """Adam through the GPU's own forward/backward kernels on a fixed target function."""
    print("\n[gpu mlp] training", flush=True)
    rng = np.random.default_rng(5)
    n = 1 << 15
    x = random_x(rng, n)
    proj = rng.normal(0, 1.0, (29, 12))
    t = np.tanh(x[:, :29] @ proj * 0.3).astype(np.float32)
    renderer.mlp_init(11)
    renderer.mlp_reserve_samples(n)
    scale = np.ones(12, np.float32)
    loss0, _, _ = renderer.mlp_gradients(x[:4096], t[:4096], scale)
    renderer.mlp_set_samples(x, t)
    last = renderer.mlp_train(epochs=30, batch=2048, lr=3e-3, scale=scale, seed=2)
    loss1, _, _ = renderer.mlp_gradients(x[:4096], t[:4096], scale)
    print(f"loss {loss0:.4f} -> {loss1:.5f} (last epoch {last:.5f})")
    assert loss1 < 0.1 * loss0

