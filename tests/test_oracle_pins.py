"""Known answers and closed forms that pin the oracle's grid sizing, its
rest-chord bend term, its max_step cap and its curvature projection.

Each check is chosen so that a plausible slip in the oracle fails it:
- grid sizing (crates/tangle_relax/src/device/world.rs:322-417): hand-made
  worlds whose sizes are exact in f32, one per term of the rule (skin from the
  largest radius, margin from max(2 max_step, skin), longest of rest and
  initial length, cell-size scale, floor of the cell count), and the two
  benchmark workloads derived from their inputs;
- rest-chord bend (kernels.rs:1099-1161): three collinear vertices whose
  intrinsic shape is bent, so the rest chord differs from the rest length sum;
- max_step cap (kernels.rs:1482-1505): a vertex whose two segment corrections
  are each under the cap in different directions while their sum exceeds it;
- curvature projection (kernels.rs:1165-1322): a free middle vertex between
  pinned ends; one iteration must bring the turning angle to the limit angle
  up to a residual quadratic in the initial excess, and reach the limit of a
  large kink at once (a projection linearised in the cosine stays short of it).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle

f32 = np.float32


def sizing(packed, **changes):
    cfg = fx.config(**changes)
    with oracle.OracleWorld(packed, cfg) as w:
        info = w.info()
    return list(info.cells), f32(info.cell_size), f32(info.neighbor_skin)


def straight(n_segments, start=(1.0, 1.0, 1.0), step=1.0, axis=0):
    pts = []
    for i in range(n_segments + 1):
        p = list(start)
        p[axis] += step * i
        pts.append(p)
    return pts


# ---- grid sizing ------------------------------------------------------------

def test_sizing_unit_case():
    # r = 0.25, unit segments, skin scale 2 -> skin 0.5; max_step 0.125 ->
    # margin 2r + max(0.25, 0.5) = 1; block = min(1, max(1, 2)) = 1 -> cell 2;
    # a 10-long box holds 5 cells.
    packed = fx.packed([straight(4)], 0.25, [10.0] * 3, [True] * 3)
    cells, cell, skin = sizing(packed, max_step=0.125)
    assert (cells, cell, skin) == ([5, 5, 5], f32(2.0), f32(0.5))


def test_sizing_margin_takes_the_larger_of_step_and_skin():
    # max_step 0.5: 2 max_step = 1 beats the skin 0.5 -> margin 1.5, cell 2.5 -> 4 cells
    packed = fx.packed([straight(4)], 0.25, [10.0] * 3, [True] * 3)
    cells, cell, skin = sizing(packed, max_step=0.5)
    assert (cells, cell, skin) == ([4, 4, 4], f32(2.5), f32(0.5))


def test_sizing_skin_from_the_largest_radius_for_every_pair():
    # radii 0.25 and 0.5: skin 2 x 0.5 = 1, margin 2 x 0.5 + 1 = 2, cell 3 -> 3 cells
    packed = fx.packed([straight(4), straight(4, start=(1.0, 5.0, 5.0))], [0.25, 0.5], [10.0] * 3, [True] * 3)
    cells, cell, skin = sizing(packed, max_step=0.125)
    assert (cells, cell, skin) == ([3, 3, 3], f32(3.0), f32(1.0))


def test_sizing_uses_the_longer_of_rest_and_initial_length():
    # placed 2 apart with rest length 1: the initial length 2 sets the block and cell
    rest = [straight(4)]
    placed = [straight(4, step=2.0)]
    packed = fx.packed(placed, 0.25, [12.0] * 3, [True] * 3, rest=rest)
    cells, cell, skin = sizing(packed, max_step=0.125)
    assert (cells, cell) == ([4, 4, 4], f32(3.0))
    # and the rest length when it is the longer one
    packed = fx.packed(rest, 0.25, [12.0] * 3, [True] * 3, rest=placed)
    assert sizing(packed, max_step=0.125)[:2] == ([4, 4, 4], f32(3.0))


def test_sizing_scale_skin_zero_and_floor():
    packed = fx.packed([straight(4)], 0.25, [10.0, 9.0, 5.99], [True] * 3)
    # cell scale 1.5: cell 3 -> floor(10/3) = 3, floor(9/3) = 3, floor(5.99/3) = 1
    cells, cell, _ = sizing(packed, max_step=0.125, cell_size_scale=1.5)
    assert (cells, cell) == ([3, 3, 1], f32(3.0))
    # skin scale 0: margin 2r + 2 max_step = 0.75, cell 1.75
    cells, cell, skin = sizing(packed, max_step=0.125, neighbor_skin_scale=0.0)
    assert (cells, cell, skin) == ([5, 5, 3], f32(1.75), f32(0.0))
    # a box smaller than one cell still has one cell per axis
    small = fx.packed([straight(1, step=0.5)], 0.25, [1.0] * 3, [False] * 3)
    assert sizing(small, max_step=0.125)[0] == [1, 1, 1]


def f32_lengths(packed):
    pos = packed.placed.reshape(-1, 3).astype(np.float32)
    rest = []
    init = []
    start = 0
    for n in packed.counts:
        for l in range(int(n) - 1):
            a, b = start + l, start + l + 1
            d = pos[b] - pos[a]
            init.append(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
            r = packed.intrinsic[3 * b:3 * b + 3] - packed.intrinsic[3 * a:3 * a + 3]
            rest.append(f32(math.sqrt(float(r @ r))))
        start += int(n)
    return f32(max(rest)), f32(max(init))


@pytest.mark.parametrize("which", ["config1", "bimodal"])
def test_sizing_of_the_benchmark_workloads(which):
    if which == "config1":
        recipe, d, r_max = fx.config1_recipe(), 10e-6, f32(5e-6)
    else:
        recipe, d, r_max = fx.bimodal_recipe(), 7e-6, f32(9.5e-6)
    packed = recipe._packed(False)
    cfg = fx.hard_gate_settings(d).to_config(d)
    with oracle.OracleWorld(packed, cfg) as w:
        info = w.info()
    max_rest, max_init = f32_lengths(packed)
    max_len = max(max_rest, max_init)
    skin = f32(2.0) * r_max
    margin = f32(2.0) * r_max + max(f32(2.0) * f32(cfg.max_step), skin)
    cell = max_len + margin
    side = f32(packed.struct.cell_lengths[0])
    assert f32(info.neighbor_skin) == skin
    assert f32(info.cell_size) == cell
    assert list(info.cells) == [int(np.floor(side / cell))] * 3
    if which == "config1":
        # skin 10 um, margin 20 um; the f32 initial segment length is a few
        # ulps over 10 um, so the cell is just over 30 um and 300 um holds 9
        # cells per axis, not 10
        assert skin == f32(1e-5) and margin == f32(2e-5)
        assert max_init > f32(1e-5) and list(info.cells) == [9, 9, 9]
    else:
        # skin 19 um, margin 38 um, cell ~76.8 um (the longest wavy 19 um
        # segment is ~38.8 um) -> 13 cells per axis
        assert skin == f32(1.9e-5) and margin == f32(3.8e-5)
        assert list(info.cells) == [13, 13, 13]
        assert abs(float(cell) - 76.79e-6) < 0.01e-6


# ---- rest-chord bend ---------------------------------------------------------

@pytest.mark.parametrize("k_b", [0.03, 0.5, 1.0])
def test_rest_chord_bend_closed_form(k_b):
    # Intrinsic: ends 1 apart with the middle lifted 0.3 (rest lengths 0.583,
    # rest chord 1). Placed: collinear, ends 1.2 apart, middle pinned, no
    # stretch. Each Jacobi pass moves each end 0.5 k_b (L - c) toward the
    # other, so after one iteration of two passes L = c + 0.2 (1 - k_b)^2.
    rest = [[[0.0, 0.0, 0.0], [0.5, 0.3, 0.0], [1.0, 0.0, 0.0]]]
    placed = [[[1.0, 2.0, 2.0], [1.6, 2.0, 2.0], [2.2, 2.0, 2.0]]]
    pinned = np.array([0, 1, 0], dtype=np.uint8)
    packed = fx.packed(placed, 0.01, [4.0] * 3, [False] * 3, rest=rest, vertex_pinned=pinned)
    cfg = fx.config(penetration_tolerance=0.0, force_full_iterations=1, max_step=0.5, stretch_stiffness=0.0,
                    bend_stiffness=k_b)
    with oracle.OracleWorld(packed, cfg) as w:
        w.run_batch(cfg, 1)
        p = w.download_positions().reshape(-1, 3).astype(np.float64)
    assert abs((p[2, 0] - p[0, 0]) - (1.0 + 0.2 * (1.0 - k_b) ** 2)) < 1e-6
    assert abs((p[0, 0] + p[2, 0]) / 2 - 1.6) < 1e-6
    assert p[1].tolist() == np.float32([1.6, 2.0, 2.0]).astype(np.float64).tolist()
    assert np.all(np.abs(p[:, 1:] - 2.0) < 1e-7)


# ---- max_step cap --------------------------------------------------------------

def test_max_step_caps_the_sum_of_a_vertex_corrections():
    # Fiber A along x with its middle vertex at (2, 2, 2). Fiber B crosses A's
    # first segment at x = 1.5 from above (+z), fiber C its second segment at
    # x = 2.5 from the side (+y), both 0.1 apart with radii 0.1: each contact
    # has s = t = 1/2, so each endpoint of the crossed segment gets cf pen / 2
    # = 0.04 along the normal (contact.rs:834-863). The middle vertex gets
    # (0, -0.04, -0.04), |.| = 0.0566; with max_step 0.05 each part is under
    # the cap and the sum is scaled to 0.05 along (0, -1, -1) / sqrt 2.
    a = [[1.0, 2.0, 2.0], [2.0, 2.0, 2.0], [3.0, 2.0, 2.0]]
    b = [[1.5, 1.0, 2.1], [1.5, 3.0, 2.1]]
    c = [[2.5, 2.1, 1.0], [2.5, 2.1, 3.0]]
    packed = fx.packed([a, b, c], 0.1, [4.0] * 3, [False] * 3)
    cfg = fx.config(penetration_tolerance=0.0, force_full_iterations=1, max_step=0.05)
    with oracle.OracleWorld(packed, cfg) as w:
        s = w.run_batch(cfg, 1)
        p = w.download_positions().reshape(-1, 3).astype(np.float64)
    move = p[1] - np.array([2.0, 2.0, 2.0])
    assert abs(np.linalg.norm(move) - 0.05) < 1e-7
    assert abs(move[0]) < 1e-7 and abs(move[1] - move[2]) < 1e-7 and move[1] < 0
    # the ends carry one correction each, uncapped
    assert abs(p[0, 2] - (2.0 - 0.04)) < 1e-7 and abs(p[0, 1] - 2.0) < 1e-7
    assert abs(p[2, 1] - (2.0 - 0.04)) < 1e-7 and abs(p[2, 2] - 2.0) < 1e-7
    assert abs(s.max_displacement - 0.05) < 1e-7


# ---- curvature projection -------------------------------------------------------

def turning_angle(p):
    u, v = p[1] - p[0], p[2] - p[1]
    c = float(u @ v) / (np.linalg.norm(u) * np.linalg.norm(v))
    return math.acos(max(-1.0, min(1.0, c))), float(np.linalg.norm(u) + np.linalg.norm(v))


def projected_excess(theta0, kappa_max, margin=1e-3):
    """Kink of turning angle theta0 between pinned ends 1 apart; returns the
    excess over the limit angle before and after one iteration (one forward
    and one backward projection of the only triplet)."""
    h = 0.5 * math.tan(theta0 / 2)
    placed = [[[1.0, 1.0, 1.0], [1.5, 1.0 + h, 1.0], [2.0, 1.0, 1.0]]]
    packed = fx.packed(placed, 0.001, [4.0] * 3, [False] * 3, pin_fiber_ends=True, bend_radius=1.0 / kappa_max)
    cfg = fx.config(penetration_tolerance=1.0, force_full_iterations=1, max_step=0.5, stretch_stiffness=0.0,
                    bend_stiffness=0.0, constraint_iterations=1, curvature_cleanup_sweeps=1,
                    curvature_limit_safety_margin=margin, curvature_limit_stiffness=1.0)
    kappa = kappa_max * (1.0 - margin)

    def limit(length_sum):
        return 2.0 * math.asin(0.25 * kappa * length_sum)

    theta, lsum = turning_angle(np.array(placed[0]))
    before = theta - limit(lsum)
    with oracle.OracleWorld(packed, cfg) as w:
        w.run_batch(cfg, 1)
        p = w.download_positions().reshape(-1, 3).astype(np.float64)
    assert p[0].tolist() == np.float32(placed[0][0]).astype(np.float64).tolist()
    assert p[2].tolist() == np.float32(placed[0][2]).astype(np.float64).tolist()
    theta, lsum = turning_angle(p)
    return before, theta - limit(lsum)


def test_projection_removes_the_excess_angle_to_second_order():
    # limit ~0.2 rad; after one iteration the remaining excess is a small
    # multiple of the squared initial excess (a cosine-linearised projection
    # leaves 3.9e-4 at excess 0.078 and 6.1e-5 at 0.038)
    for excess in (0.08, 0.04, 0.02):
        before, after = projected_excess(0.2 + excess, 0.4)
        assert before > 0.5 * excess
        assert abs(after) <= 0.005 * before ** 2 + 5e-6, (before, after)


@pytest.mark.parametrize("kink", [0.6, 0.8])
def test_projection_removes_a_large_kink_at_once(kink):
    # limit ~0.1 rad: one iteration ends at or just past the limit (the
    # angle form overshoots by a second-order amount); a projection
    # linearised in the cosine moves the angle by about (theta^2 -
    # theta_min^2) / (2 theta) per projection and is still 12-13 % of the
    # excess short of the limit after two
    before, after = projected_excess(kink, 0.2)
    assert before > 0.45
    assert -0.15 * before <= after <= 0.0, (before, after)


# ---------------------------------------------------------------------------
# Curvature cleanup counts per fiber (the work of the GPU's sequential
# sweep): the per-fiber projections sum to the oracle's own counter of the
# same call, every applied projection is a visited triplet, and a fiber is
# visited at least once forward. On the benchmark felt after 200 iterations
# the test prints the histogram of projections per fiber (run with -s).


def curvature_counts(scale, iterations):
    import ctypes as C
    from tangle_amd import RelaxationSettings

    packed = fx.bimodal_recipe(scale=scale)._packed(False)
    cfg = RelaxationSettings().to_config(7e-6)
    cfg.force_full_iterations = 1
    counters = (C.c_uint64 * 4)()
    with oracle.OracleWorld(packed, cfg) as w:
        w.run_batch(cfg, iterations)
        oracle.library().orc_gs_counters(counters)  # clears them
        w.run_batch(cfg, 1)  # one iteration: one cleanup (the last pass only measures)
        oracle.library().orc_gs_counters(counters)
        projections, visits = w.curvature_fiber_counts()
    return packed, list(counters), projections.astype(np.int64), visits.astype(np.int64)


@pytest.mark.parametrize("scale,iterations", [(0.02, 20), (1.0, 200)])
def test_curvature_fiber_counts(scale, iterations):
    packed, counters, projections, visits = curvature_counts(scale, iterations)
    calls, total, _, most = counters
    assert calls == 1
    assert int(projections.sum()) == total and int(projections.max()) == most
    inner = np.maximum(packed.counts.astype(np.int64) - 2, 0)
    assert (projections <= visits).all() and (visits >= inner).all()
    if scale == 1.0:
        edges = [0, 1, 2, 5, 10, 20, 50, 100, 200, 500, 1000]
        hist, _ = np.histogram(projections, bins=edges + [1 << 30])
        print(f"\n[curvature] fibers {projections.size}, projections {total}, visits {int(visits.sum())}, "
              f"largest per fiber {most} projections / {int(visits.max())} visits")
        for lo, hi, n in zip(edges, edges[1:] + [None], hist):
            print(f"[curvature] projections {lo}..{'' if hi is None else hi - 1}: {n} fibers")
