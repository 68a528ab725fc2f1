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


# ---------------------------------------------------------------------------
# Segment features (k_features in tangle_amd/csrc/learned.hip), float32
# arithmetic in the kernel's order; the closest points follow the capsule
# pair routine of the contact pass (contact/geometry.rs closest points of two
# segments, minimum image per axis).

def _f(x):
    return np.float32(x)


def closest_offset(p1, q1, p2, q2, periodic, lengths):
    from .cell_hash import rust_round
    eps_len, eps_par = _f(1.0e-20), _f(1.0e-6)
    p1 = [_f(v) for v in p1]
    q1 = [_f(v) for v in q1]
    p2 = [_f(v) for v in p2]
    q2 = [_f(v) for v in q2]
    half = _f(0.5)
    for k in range(3):
        mid = half * (((p2[k] + q2[k]) - p1[k]) - q1[k])
        if periodic[k]:
            sh = _f(rust_round(np.float32(mid / _f(lengths[k])))) * _f(lengths[k])
            p2[k] = _f(p2[k] - sh)
            q2[k] = _f(q2[k] - sh)
    d1 = [q1[k] - p1[k] for k in range(3)]
    d2 = [q2[k] - p2[k] for k in range(3)]
    r = [p1[k] - p2[k] for k in range(3)]

    def dot(u, v):
        return _f(_f(_f(u[0] * v[0]) + _f(u[1] * v[1])) + _f(u[2] * v[2]))

    def clamp01(x):
        return _f(min(max(x, _f(0.0)), _f(1.0)))

    a, e, f = dot(d1, d1), dot(d2, d2), dot(d2, r)
    s = t = _f(0.0)
    if a <= eps_len and e > eps_len:
        t = clamp01(f / e)
    elif a > eps_len:
        c = dot(d1, r)
        if e <= eps_len:
            s = clamp01(-c / a)
        else:
            b = dot(d1, d2)
            den = _f(a * e) - _f(b * b)
            if abs(den) > _f(_f(eps_par * a) * e):
                s = clamp01(_f(_f(b * f) - _f(c * e)) / den)
            proj = _f(_f(b * s) + f) / e
            if proj < 0:
                s = clamp01(-c / a)
            elif proj > 1:
                t = _f(1.0)
                s = clamp01(_f(b - c) / a)
            else:
                t = proj
    o = [_f(_f(p2[k] + _f(d2[k] * t)) - _f(p1[k] + _f(d1[k] * s))) for k in range(3)]
    dist = _f(np.sqrt(dot(o, o)))
    return o, dist


def segment_features(positions, seg_v0, seg_radius, lists, counts, periodic, lengths, neighbors=16, history=None):
    """Feature rows (bf16 bits, 64 per segment) from per-segment candidate
    lists: the `neighbors` candidates of smallest surface gap (earlier list
    entries first on ties), each as unit direction times
    min((r1 + r2) / distance, 8); the unit tangent; then, from `history`
    (segment midpoints at the last mark, S x 3, or None), h = (midpoint -
    history) / diameter, m = |h|: h / m and log10(m) / 8 + 1 (zero when
    m <= 1e-12 or no history); then zeros."""
    pos = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    S = len(seg_v0)
    out = np.zeros((S, 64), dtype=np.float32)
    for s in range(S):
        v = int(seg_v0[s])
        p1, q1 = pos[v], pos[v + 1]
        r1 = _f(seg_radius[s])
        cands = []
        for k in range(int(counts[s])):
            o_seg = int(lists[s][k])
            w = int(seg_v0[o_seg])
            o, dist = closest_offset(p1, q1, pos[w], pos[w + 1], periodic, lengths)
            gap = _f(dist - _f(r1 + _f(seg_radius[o_seg])))
            cands.append((gap, k, o, dist, _f(r1 + _f(seg_radius[o_seg]))))
        cands.sort(key=lambda c: (c[0], c[1]))
        for j, (gap, _, o, dist, contact) in enumerate(cands[:neighbors]):
            d = _f(max(dist, _f(_f(1.0e-3) * contact)))
            w = _f(_f(min(_f(contact / d), _f(8.0))) / d)
            out[s, 3 * j:3 * j + 3] = [_f(o[0] * w), _f(o[1] * w), _f(o[2] * w)]
        t = [_f(q1[k] - p1[k]) for k in range(3)]
        n = _f(np.sqrt(_f(_f(_f(t[0] * t[0]) + _f(t[1] * t[1])) + _f(t[2] * t[2]))))
        inv = _f(_f(1.0) / _f(max(n, _f(1.0e-30))))
        out[s, 48:51] = [_f(t[0] * inv), _f(t[1] * inv), _f(t[2] * inv)]
        if history is not None:
            hist = np.asarray(history, dtype=np.float32).reshape(-1, 3)[s]
            diameter = _f(_f(2.0) * r1)
            h = [_f(_f(_f(_f(0.5) * _f(p1[k] + q1[k])) - _f(hist[k])) / diameter) for k in range(3)]
            m = _f(np.sqrt(_f(_f(_f(h[0] * h[0]) + _f(h[1] * h[1])) + _f(h[2] * h[2]))))
            if m > _f(1.0e-12):
                out[s, 51:54] = [_f(h[0] / m), _f(h[1] / m), _f(h[2] / m)]
                out[s, 54] = _f(_f(np.log10(m)) * _f(0.125)) + _f(1.0)
    return f32_to_bf16_bits(out)
