"""ORACLE — test infrastructure only.

numpy restatement of the learned warm start's network (tangle_amd/csrc/learned.hip):
y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 with weights [out][in] stored as
one f32 parameter vector (W1 64x64, b1 64, W2 64x64, b2 64, W3 16x64, b3 16),
3 outputs used, loss = mean over rows of the squared error summed over the
3 outputs. `bf16=True` rounds the matrix-core operands (weights, activations,
output and hidden gradients) to bfloat16 at the points the kernels do and
accumulates in float64, which the kernels' FP32 accumulation must match
within rounding; `bf16=False` is the plain float64 network, whose gradients
are pinned by central finite differences.
"""

from __future__ import annotations

import numpy as np

OFF_W1, OFF_B1 = 0, 4096
OFF_W2, OFF_B2 = 4160, 8256
OFF_W3, OFF_B3 = 8320, 9344
PARAMS = 9360


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(64, 64), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(16, 64), p[OFF_B3:PARAMS])


def _q(a, on):
    return bf16_round(np.asarray(a, dtype=np.float32)).astype(np.float64) if on else np.asarray(a, np.float64)


def forward(params, x, bf16=True):
    w1, b1, w2, b2, w3, b3 = unpack(params)
    w1, w2, w3 = _q(w1, bf16), _q(w2, bf16), _q(w3, bf16)
    x = np.asarray(x, np.float64)
    h1 = _q(np.maximum(x @ w1.T + b1, 0.0), bf16)
    h2 = _q(np.maximum(h1 @ w2.T + b2, 0.0), bf16)
    y = h2 @ w3.T + b3
    return y[:, :3], (x, h1, h2, w1, w2, w3)


def loss_and_gradients(params, x, t, bf16=True):
    rows = x.shape[0]
    y, (x, h1, h2, w1, w2, w3) = forward(params, x, bf16)
    e = y - np.asarray(t, np.float64)
    loss = float(np.sum(e * e) / rows)
    dy = np.zeros((rows, 16))
    dy[:, :3] = _q(2.0 * e / rows, bf16)
    g = np.zeros(PARAMS)
    g[OFF_W3:OFF_B3] = (dy.T @ h2).reshape(-1)
    g[OFF_B3:PARAMS] = dy.sum(axis=0)
    dz2 = _q((dy @ w3) * (h2 > 0), bf16)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = _q((dz2 @ w2) * (h1 > 0), bf16)
    g[OFF_W1:OFF_B1] = (dz1.T @ x).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, g


def finite_difference(params, x, t, index, h=1e-6):
    p = np.asarray(params, np.float64).copy()
    p[index] += h
    lp, _ = loss_and_gradients(p, x, t, bf16=False)
    p[index] -= 2 * h
    lm, _ = loss_and_gradients(p, x, t, bf16=False)
    return (lp - lm) / (2 * h)
