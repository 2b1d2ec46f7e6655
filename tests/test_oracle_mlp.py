"""Pins of the MLP oracle (oracle/mlp.py) by what the mathematics fixes: the
float64 gradients against central finite differences of the loss, and the
bfloat16 rounding against hand-worked values (round to nearest, ties to even)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5, 3.0e38,
                  1.0 + 2.0 ** -8 + 2.0 ** -20], np.float32)
    want = [1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 2.9980e38, 1.0 + 2.0 ** -7]
    got = M.bf16(x)
    # 1 + 2^-8 is a tie: to even (1.0); 1 + 2^-7 + 2^-8 is a tie: to even (1 + 2^-6);
    # just above a tie rounds up.
    for g, w_ in zip(got[[0, 1, 2, 3, 4, 6]], [want[i] for i in (0, 1, 2, 3, 4, 6)]):
        assert g == w_
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2.0 ** -8


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_gradients_match_finite_differences(layer):
    rng = np.random.default_rng(10 + layer)
    ws, bs = M.init(seed=layer)
    ws = [w.astype(np.float64) for w in ws]
    bs = [rng.normal(0, 0.1, b.shape) for b in bs]
    if M.N_OUT < 16:
        bs[3][M.N_OUT:] = 0.0
    x = np.zeros((64, 32))
    x[:, :M.N_IN] = rng.uniform(-1, 1, (64, M.N_IN))
    t = rng.normal(0, 0.5, (64, M.N_OUT))
    scale = rng.uniform(0.5, 2.0, M.N_OUT)
    _, dws, dbs = M.loss_and_grads(ws, bs, x, t, scale, bf16_operands=False)
    h = 1e-6
    rows = rng.integers(0, ws[layer].shape[0] if layer < 3 else M.N_OUT, 6)
    cols = rng.integers(0, ws[layer].shape[1] if layer else M.N_IN, 6)
    for r, c in zip(rows, cols):
        wp = [w.copy() for w in ws]
        wm = [w.copy() for w in ws]
        wp[layer][r, c] += h
        wm[layer][r, c] -= h
        lp, _, _ = M.loss_and_grads(wp, bs, x, t, scale, bf16_operands=False)
        lm, _, _ = M.loss_and_grads(wm, bs, x, t, scale, bf16_operands=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - dws[layer][r, c]) <= 1e-6 + 1e-5 * abs(fd), (layer, r, c, fd, dws[layer][r, c])
    for r in rows:
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[layer][r] += h
        bm[layer][r] -= h
        lp, _, _ = M.loss_and_grads(ws, bp, x, t, scale, bf16_operands=False)
        lm, _, _ = M.loss_and_grads(ws, bm, x, t, scale, bf16_operands=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - dbs[layer][r]) <= 1e-6 + 1e-5 * abs(fd)
