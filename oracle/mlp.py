"""THE ORACLE for the learned shading term (test infrastructure).

A float64 restatement of the surrogate network the HIP kernels train and run:
  x (32 features) -> 64 -> 64 -> 64 (ReLU) -> 16 outputs.
Weights are stored out x in (row-major), biases per output. With
`bf16_operands=True` the operands are rounded to bfloat16 (round to nearest
even) exactly where the kernels feed the matrix cores - weights, inputs and
every hidden activation, and the backward pass's output errors - while
accumulation stays in float64, which
isolates the kernels' float32 accumulation error.

The loss is the mean over samples and the N_OUT used outputs of
((y - t) * s)^2, with per-output scales s.
"""
from __future__ import annotations

import numpy as np

DIMS = (32, 64, 64, 64, 16)
N_OUT = 16
N_IN = 32  # features used (inputs past N_IN are zero-padded)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    rounded = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def init(seed: int = 0):
    """He-uniform weights, zero biases (the kernels' default initialisation)."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for i in range(4):
        fan_in = DIMS[i] if i else N_IN
        lim = np.sqrt(6.0 / fan_in)
        w = rng.uniform(-lim, lim, (DIMS[i + 1], DIMS[i]))
        if i == 0 and N_IN < DIMS[0]:
            w[:, N_IN:] = 0.0
        if i == 3 and N_OUT < DIMS[4]:
            w[N_OUT:, :] = 0.0
        ws.append(w.astype(np.float32))
        bs.append(np.zeros(DIMS[i + 1], np.float32))
    return ws, bs


def forward(ws, bs, x, bf16_operands=True):
    """Returns (y, cache). x: [B, 32]."""
    q = bf16 if bf16_operands else (lambda a: np.asarray(a, np.float64))
    h = q(x)
    acts, pre = [h], []
    for i in range(4):
        z = h @ q(ws[i]).T + np.asarray(bs[i], np.float64)
        pre.append(z)
        if i < 3:
            a = np.maximum(z, 0.0)
            h = q(a)
            acts.append(h)
        else:
            h = z
    return h, (acts, pre)


def loss_and_grads(ws, bs, x, t, scale, bf16_operands=True):
    """Loss and float64 gradients dW, db for one batch (see module docstring)."""
    y, (acts, pre) = forward(ws, bs, x, bf16_operands)
    B = x.shape[0]
    s = np.zeros(DIMS[-1])
    s[:N_OUT] = scale
    diff = (y - np.pad(t, ((0, 0), (0, DIMS[-1] - t.shape[1])))) * s
    loss = float((diff[:, :N_OUT] ** 2).sum() / (B * N_OUT))
    dz = 2.0 * diff * s / (B * N_OUT)
    q = bf16 if bf16_operands else (lambda a: np.asarray(a, np.float64))
    dws, dbs = [None] * 4, [None] * 4
    for i in (3, 2, 1, 0):
        dws[i] = q(dz).T @ acts[i]
        dbs[i] = dz.sum(0)
        if i:
            dh = q(dz) @ q(ws[i])
            dz = dh * (pre[i - 1] > 0)
    return loss, dws, dbs
