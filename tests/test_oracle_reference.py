"""Pins the CPU oracle (oracle/relax_oracle.cpp) before it is trusted.

Checks: the reference's own device-world test expectations
(crates/tangle_relax/src/device/world_tests.rs), closed forms for the stretch,
contact and curvature terms, the neighbour-list completeness invariant with
an f64 host distance (world_tests.rs:1365-1485), and an independent numpy
float32 restatement of the cell hash and of the list predicate evaluated by
brute force over all pairs (oracle/cell_hash.py), which must match bit for bit.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle
from oracle import cell_hash


def run(packed, iterations, **changes):
    cfg = fx.config(**changes)
    w = oracle.OracleWorld(packed, cfg)
    s = w.run_batch(cfg, iterations)
    return w, s


@pytest.mark.parametrize("case", fx.REFERENCE_CASES, ids=[c[0] for c in fx.REFERENCE_CASES])
def test_reference_expectations(case):
    name, builder, changes, iterations, expect, _ = case
    w, s = run(builder(), iterations, **changes)
    assert expect(s), (name, s.as_dict())
    w.close()


def test_curvature_limit_moves_free_middle_between_pinned_ends():
    # world_tests.rs:461-509: pinned ends keep their exact f32 coordinates
    w, s = run(fx.bent_triplet(), 40, penetration_tolerance=1e-5, stretch_stiffness=0.0, bend_stiffness=0.0,
               constraint_iterations=1, max_step=0.02)
    p = w.download_positions()
    assert p[0:3].tolist() == np.float32([0.8, 1.0, 1.0]).tolist()
    assert p[6:9].tolist() == np.float32([1.2, 1.0, 1.0]).tolist()
    assert p[4] < np.float32(1.15)


def test_zigzag_converges_in_budget_with_exact_ends():
    # world_tests.rs:1159-1188 (budget 30; 22 on the reference's Metal run)
    w, s = run(fx.zigzag(), 30, penetration_tolerance=1e-5, stretch_stiffness=0.0, bend_stiffness=0.0,
               constraint_iterations=1, curvature_cleanup_sweeps=1, max_step=0.02)
    assert s.converged and s.max_curvature_ratio <= 1 + 1e-5
    assert s.total_iterations <= 30
    p = w.download_positions()
    assert p[0:3].tolist() == np.float32([0.2, 1.0, 1.0]).tolist()
    assert p[240:243].tolist() == np.float32([1.8, 1.0, 1.0]).tolist()


@pytest.mark.parametrize("builder,gap,tol", [(fx.periodic_pair, 0.1, 1e-5), (fx.si_crossing, 20e-6, 1e-8),
                                             (fx.unwrapped_images, 20e-6, 1e-8)])
def test_initial_overlap_known_answers(builder, gap, tol):
    # world_tests.rs:953-956, 992-995, 1022-1025 (surface gap -0.1, -20e-6);
    # a tolerance every state meets stops the batch at its first pass.
    w, s = run(builder(), 1, penetration_tolerance=1.0, max_step=1e-6)
    assert s.total_iterations == 0
    assert abs(s.max_penetration - gap) < tol


def test_self_image_contact_is_pushed_apart():
    # world_tests.rs:1029-1064
    cfg = fx.config(penetration_tolerance=1e-4, max_step=0.05)
    with oracle.OracleWorld(fx.self_image(), cfg) as w:
        assert w.run_batch(cfg, 1).max_penetration > 0.05
        assert w.run_batch(cfg, 999).max_penetration < 0.005


def test_listed_overflow_and_skin0_paths_agree():
    # world_tests.rs:1309-1337
    def relax(**cell):
        w, _ = run(fx.dense_mat(), 40, penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01, **cell)
        return w.download_positions(), w.info().neighbor_list_rebuilds

    every, every_rebuilds = relax(neighbor_skin_scale=0.0)
    listed, listed_rebuilds = relax()
    overflowed, _ = relax(neighbor_capacity=1)
    assert np.max(np.abs(listed - every)) < 1e-4
    assert np.max(np.abs(overflowed - every)) < 1e-4
    assert listed_rebuilds < every_rebuilds


def test_bitwise_repeatable():
    # world_tests.rs:1487-1510
    for cap in (48, 1):
        a, _ = run(fx.dense_mat(), 40, penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01,
                   neighbor_capacity=cap)
        b, _ = run(fx.dense_mat(), 40, penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01,
                   neighbor_capacity=cap)
        assert np.array_equal(a.download_positions().view(np.uint32), b.download_positions().view(np.uint32))
        assert a.info().neighbor_list_rebuilds == b.info().neighbor_list_rebuilds


def host_segment_axis_distance(p1, q1, p2, q2):
    """f64 closest distance of two segment axes (world_tests.rs:1365-1398)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f, c, b = d1 @ d1, d2 @ d2, d2 @ r, d1 @ r, d1 @ d2
    den = a * e - b * b
    s = min(max((b * f - c * e) / den, 0.0), 1.0) if den > 1e-12 * a * e else 0.0
    t = (b * s + f) / e
    if t < 0.0:
        t, s = 0.0, min(max(-c / a, 0.0), 1.0)
    elif t > 1.0:
        t, s = 1.0, min(max((b - c) / a, 0.0), 1.0)
    delta = p2 + d2 * t - p1 - d1 * s
    return math.sqrt(delta @ delta)


def test_neighbor_lists_hold_every_pair_within_the_skin():
    # world_tests.rs:1399-1485
    packed = fx.dense_mat()
    w, _ = run(packed, 1, max_step=0.01)
    counts, lists = w.neighbor_lists()
    info = w.info()
    skin = float(info.neighbor_skin)
    pos = packed.placed.reshape(-1, 3).astype(np.float32).astype(np.float64)
    seg = segments_of(packed)
    lengths, periodic = [2.4, 2.4, 1.0], [True, True, False]
    checked = 0
    for first in range(len(seg)):
        assert counts[first] <= info.neighbor_block, f"segment {first} overflowed"
        listed = lists[first, :counts[first]]
        assert len(set(listed.tolist())) == len(listed)
        a, b, fa = seg[first]
        p1, q1 = pos[a], pos[b]
        for second in range(len(seg)):
            c, d, fb = seg[second]
            adjacent = fa == fb and len({a, b} & {c, d}) > 0
            if second == first or adjacent:
                assert second not in listed
                continue
            p2, q2 = pos[c].copy(), pos[d].copy()
            for k in range(3):
                if periodic[k]:
                    shift = round_half_away(0.5 * (p2[k] + q2[k] - p1[k] - q1[k]) / lengths[k]) * lengths[k]
                    p2[k] -= shift
                    q2[k] -= shift
            interaction = 0.2 + skin
            distance = host_segment_axis_distance(p1, q1, p2, q2)
            if distance < 0.999 * interaction:
                assert second in listed, (first, second, distance)
                checked += 1
            elif distance > 1.001 * interaction:
                assert second not in listed
    assert checked > len(seg)


def round_half_away(x):
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def segments_of(packed):
    out, start = [], 0
    for f, n in enumerate(packed.counts):
        for l in range(int(n) - 1):
            out.append((start + l, start + l + 1, f))
        start += int(n)
    return out


def seg_arrays(packed):
    seg = segments_of(packed)
    vtx = np.array([(a, b) for a, b, _ in seg], dtype=np.int64)
    fib = np.array([f for _, _, f in seg], dtype=np.int64)
    rad = packed.radius.astype(np.float32)[fib]
    return vtx, fib, rad


def arc_lengths(packed):
    intr = packed.intrinsic.reshape(-1, 3).astype(np.float32)
    arc = np.zeros(intr.shape[0], dtype=np.float32)
    start = 0
    for n in packed.counts:
        total = 0.0
        for v in range(start + 1, start + int(n)):
            d = intr[v].astype(np.float64) - intr[v - 1].astype(np.float64)
            total += math.sqrt(float(d @ d))
            arc[v] = np.float32(total)
        start += int(n)
    return arc


@pytest.mark.parametrize("which", ["dense_mat", "config1", "bimodal_small"])
def test_cell_list_and_pairs_match_numpy_restatement(which):
    if which == "dense_mat":
        packed, cfg = fx.dense_mat(), fx.config(max_step=0.01)
    elif which == "config1":
        packed = fx.config1_recipe()._packed(False)
        cfg = fx.hard_gate_settings(10e-6).to_config(10e-6)
    else:
        packed = fx.bimodal_recipe(scale=0.01)._packed(False)
        cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    with oracle.OracleWorld(packed, cfg) as w:
        w.rebuild_lists()
        info = w.info()
        counts, offsets, segs = w.cell_list()
        pos = w.download_positions().reshape(-1, 3)
        vtx, fib, rad = seg_arrays(packed)
        lower = packed.struct.cell_origin[:]
        upper = [lower[k] + packed.struct.cell_lengths[k] for k in range(3)]
        lower32 = [np.float32(v) for v in lower]
        upper32 = [np.float32(v) for v in upper]
        periodic = [bool(p) for p in packed.struct.cell_periodic]
        rc, ro, rs = cell_hash.cell_list(pos, vtx, lower32, upper32, list(info.cells), periodic)
        assert np.array_equal(counts, rc)
        assert np.array_equal(offsets, ro)
        assert np.array_equal(segs, rs)
        ncount, lists = w.neighbor_lists()
        brute = cell_hash.brute_force_lists(pos, vtx, fib, rad, arc_lengths(packed), lower32, upper32, periodic,
                                            np.float32(info.neighbor_skin))
        for s in range(len(brute)):
            assert ncount[s] == len(brute[s]), (s, ncount[s], brute[s])
            if ncount[s] <= info.neighbor_block:
                assert np.array_equal(lists[s, :ncount[s]], brute[s])


def test_rust_round_is_half_away_from_zero():
    x = np.float32([0.5, 1.5, 2.5, -0.5, -1.5, 0.49999997, -2.5000002])
    assert cell_hash.rust_round(x).tolist() == [1.0, 2.0, 3.0, -1.0, -2.0, 0.0, -3.0]


def test_stretch_closed_form():
    # Two vertices, rest length 1 placed 1.2 apart: each Jacobi pass removes
    # the fraction k_s of the excess (0.5 k_s per end), so after one iteration
    # of two passes the length is 1 + 0.2 (1 - k_s)^2.
    packed = fx.packed([[[1.0, 2.0, 2.0], [2.2, 2.0, 2.0]]], 0.01, [4.0] * 3, [False] * 3,
                       rest=[[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]])
    w, s = run(packed, 1, penetration_tolerance=0.0, force_full_iterations=1, max_step=0.5)
    p = w.download_positions().reshape(-1, 3).astype(np.float64)
    assert abs((p[1, 0] - p[0, 0]) - (1.0 + 0.2 * 0.65 ** 2)) < 1e-6
    assert abs((p[0, 0] + p[1, 0]) / 2 - 1.6) < 1e-6


def test_contact_closed_form_crossing():
    # Perpendicular segments crossing at their midpoints 0.1 apart, radii 0.1:
    # penetration 0.1, s = t = 1/2, denominator 4 * 1/4 = 1, so each endpoint
    # moves cf * pen / 2 = 0.04 along the normal (contact.rs:834-863).
    packed = fx.packed([[[1.0, 2.0, 2.0], [3.0, 2.0, 2.0]], [[2.0, 1.0, 2.1], [2.0, 3.0, 2.1]]], 0.1, [4.0] * 3,
                       [False] * 3)
    w, s = run(packed, 1, penetration_tolerance=0.0, force_full_iterations=1, max_step=0.2)
    p = w.download_positions().reshape(-1, 3).astype(np.float64)
    assert np.allclose(p[:2, 2], 2.0 - 0.04, atol=1e-6)
    assert np.allclose(p[2:, 2], 2.1 + 0.04, atol=1e-6)
    assert abs(s.max_displacement - 0.04) < 1e-6


def test_curvature_ratio_of_circle_samples():
    # curvature.rs:41-67: circle samples of radius R give kappa = 1/R; with a
    # bend limit of radius R/2 the measured ratio is 1/2.
    R = 2.0
    n = 9
    pts = [[2.0 + R * math.cos(i * math.pi / 2 / (n - 1)), 2.0 + R * math.sin(i * math.pi / 2 / (n - 1)), 4.0]
           for i in range(n)]
    packed = fx.packed([pts], 0.01, [8.0] * 3, [False] * 3, bend_radius=R / 2)
    w, s = run(packed, 1, penetration_tolerance=1.0, curvature_ratio_tolerance=1.0, max_step=0.01)
    assert s.total_iterations == 0
    assert abs(s.max_curvature_ratio - 0.5) < 1e-5
