"""Pins the oracle of the learned warm start (oracle/mlp_ref.py) on the CPU:
bfloat16 rounding known answers, backpropagation against central finite
differences of the float64 network, and the segment features against
closed-form geometry."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38, 1.0e-40], dtype=np.float32)
    bits = mlp_ref.f32_to_bf16_bits(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even keeps 1.0;
    # 1 + 3*2^-9 lies above the tie and rounds up
    assert bits.tolist() == [0x3F80, 0x3F80, 0x3F81, 0xC020, 0x7F62, 0x0001]
    assert mlp_ref.bf16_bits_to_f32(np.array([0x3F81], np.uint16))[0] == np.float32(1.0 + 2.0 ** -7)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = rng.uniform(-0.4, 0.4, mlp_ref.PARAMS)
    x = rng.uniform(-1, 1, (7, 64))
    t = rng.uniform(-1, 1, (7, 3))
    _, g = mlp_ref.loss_and_gradients(params, x, t, bf16=False)
    picks = [mlp_ref.OFF_W1 + 5 * 64 + 9, mlp_ref.OFF_B1 + 17, mlp_ref.OFF_W2 + 63 * 64 + 1, mlp_ref.OFF_B2 + 40,
             mlp_ref.OFF_W3 + 2 * 64 + 33, mlp_ref.OFF_W3 + 0 * 64 + 7, mlp_ref.OFF_B3 + 1]
    for i in picks:
        fd = mlp_ref.finite_difference(params, x, t, i)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i
    # outputs beyond the third carry no loss, so their weights get no gradient
    assert not np.any(g[mlp_ref.OFF_W3 + 3 * 64:mlp_ref.OFF_B3])


def test_bf16_network_stays_close_to_float64():
    rng = np.random.default_rng(1)
    params = rng.uniform(-0.3, 0.3, mlp_ref.PARAMS)
    x = mlp_ref.bf16_round(rng.uniform(-1, 1, (50, 64)).astype(np.float32)).astype(np.float64)
    t = rng.uniform(-1, 1, (50, 3))
    l16, g16 = mlp_ref.loss_and_gradients(params, x, t, bf16=True)
    l64, g64 = mlp_ref.loss_and_gradients(params, x, t, bf16=False)
    assert abs(l16 - l64) <= 2e-2 * l64
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)


def test_features_closed_form():
    # segment 0 along x at y = 1; segment 1 parallel at y = 1.3; radii 0.1:
    # direction +y, ratio (0.1 + 0.1) / 0.3
    pos = np.array([[0, 1, 1], [1, 1, 1], [0, 1.3, 1], [1, 1.3, 1]], dtype=np.float32)
    v0 = np.array([0, 2])
    radius = np.array([0.1, 0.1], dtype=np.float32)
    lists = np.array([[1], [0]])
    counts = np.array([1, 1])
    f = mlp_ref.bf16_bits_to_f32(mlp_ref.segment_features(pos, v0, radius, lists, counts, [False] * 3, [4.0] * 3))
    assert f[0, 0] == 0 and f[0, 2] == 0
    assert f[0, 1] == pytest.approx(0.2 / 0.3, rel=4e-3)
    assert f[1, 1] == pytest.approx(-0.2 / 0.3, rel=4e-3)
    assert f[0, 48:51].tolist() == [1.0, 0.0, 0.0]
    assert not np.any(f[:, 3:48]) and not np.any(f[:, 51:])
    # the same pair across a periodic y boundary of length 2: self at y = 0.1,
    # other at y = 1.8; its image at y = -0.2 is the nearer one, direction -y
    pos2 = pos.copy()
    pos2[:2, 1] = np.float32(0.1)
    pos2[2:, 1] = np.float32(1.8)
    f2 = mlp_ref.bf16_bits_to_f32(
        mlp_ref.segment_features(pos2, v0, radius, lists, counts, [False, True, False], [4.0, 2.0, 4.0]))
    assert f2[0, 1] == pytest.approx(-0.2 / 0.3, rel=4e-3)


def test_features_keep_sixteen_smallest_gaps_in_order():
    # 20 parallel neighbours at increasing distances along y; list in reverse order
    n = 20
    pos = [[0, 0, 0], [1, 0, 0]]
    for k in range(n):
        y = 0.25 + 0.05 * k
        pos += [[0, y, 0], [1, y, 0]]
    pos = np.array(pos, dtype=np.float32)
    v0 = np.arange(0, 2 * (n + 1), 2)
    radius = np.full(n + 1, 0.1, dtype=np.float32)
    lists = np.zeros((n + 1, n), dtype=np.int64)
    lists[0] = np.arange(n, 0, -1)
    counts = np.zeros(n + 1, dtype=np.int64)
    counts[0] = n
    f = mlp_ref.bf16_bits_to_f32(mlp_ref.segment_features(pos, v0, radius, lists, counts, [False] * 3, [9.0] * 3))
    ys = f[0, 1:48:3]
    expect = [0.2 / (0.25 + 0.05 * k) for k in range(16)]
    assert np.allclose(ys, expect, rtol=4e-3)


def test_features_motion_history_closed_form():
    # segment 0 moved +0.002 along y since the mark; diameter 0.2: h = 0.01,
    # direction +y, log10(0.01) / 8 + 1 = 0.75; segment 1 did not move
    pos = np.array([[0, 1, 1], [1, 1, 1], [0, 1.3, 1], [1, 1.3, 1]], dtype=np.float32)
    v0 = np.array([0, 2])
    radius = np.array([0.1, 0.1], dtype=np.float32)
    mids = np.array([[0.5, 1.0 - 0.002, 1.0], [0.5, 1.3, 1.0]], dtype=np.float32)
    f = mlp_ref.bf16_bits_to_f32(mlp_ref.segment_features(
        pos, v0, radius, np.zeros((2, 0), dtype=np.int64), np.zeros(2, dtype=np.int64), [False] * 3, [4.0] * 3,
        history=mids))
    assert f[0, 51] == 0 and f[0, 53] == 0
    assert f[0, 52] == pytest.approx(1.0, rel=4e-3)
    assert f[0, 54] == pytest.approx(0.75, abs=4e-3)
    assert not np.any(f[1, 51:55])
