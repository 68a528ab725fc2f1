"""Pins the oracle's compaction and contact-capture restatement
(oracle/relax_oracle.cpp: compact_cell, compaction_metrics, capture_contacts)
before it is trusted as the checker of the HIP path.

Checks: the reference's own device-world expectations
(crates/tangle_relax/src/device/world_tests.rs:877-1026: rigid and affine
compaction, three capture cases), closed forms for each kinematics, for the
wall reaction and face pressure of a moving wall, and for the stretch,
contact and bending energies, and a brute-force all-pairs capture in float64
with the minimum image, which the cell-grid traversal must reproduce.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle

RIGID, WALLS, AFFINE = 0, 1, 2


def single_fiber():
    # world_tests.rs:878-887: one 1-long fiber in a 4-wide bounded cell
    return fx.packed([[[1.0, 1.0, 1.0], [2.0, 1.0, 1.0]]], 0.1, [4.0] * 3, [False] * 3)


def test_rigid_center_compaction_keeps_length_and_moves_center():
    # world_tests.rs:876-904
    with oracle.OracleWorld(single_fiber(), fx.config(max_step=0.1)) as w:
        w.compact_cell([0.0, 0.0, 0.0], [2.0, 4.0, 4.0], RIGID)
        assert w.cell_bounds() == ((0.0, 0.0, 0.0), (2.0, 4.0, 4.0))
        p = w.download_positions()
        assert abs(p[3] - p[0] - 1.0) < 1e-6
        assert abs(0.5 * (p[0] + p[3]) - 0.75) < 1e-6
        # y and z keep their fractions of unchanged axes
        assert p[1] == p[4] == np.float32(1.0) and p[2] == p[5] == np.float32(1.0)


def test_affine_compaction_scales_centerline():
    # world_tests.rs:906-932
    with oracle.OracleWorld(single_fiber(), fx.config(max_step=0.1)) as w:
        w.compact_cell([0.0, 0.0, 0.0], [2.0, 4.0, 4.0], AFFINE)
        p = w.download_positions()
        assert abs(p[3] - p[0] - 0.5) < 1e-6
        assert p[0] == np.float32(0.5) and p[3] == np.float32(1.0)


def test_moving_wall_projects_and_records_the_reaction():
    """A vertex beyond the new upper x wall is projected to wall - radius;
    the reaction is the projection, which becomes the upper-x face pressure
    k * reaction / (fraction * face area); the lower face carries none."""
    with oracle.OracleWorld(single_fiber(), fx.config(max_step=0.1)) as w:
        w.compact_cell([0.0, 0.0, 0.0], [1.5, 4.0, 4.0], WALLS)
        p = w.download_positions()
        assert p[0] == np.float32(1.0)
        assert p[3] == np.float32(np.float32(1.5) - np.float32(0.1))
        m = w.compaction_metrics(0.5, contact_stiffness=2.0, stretch_stiffness=0.0, bending_stiffness=0.0)
        reaction = np.float32(2.0) - p[3]
        expect = 2.0 * float(reaction) / 0.5 / 16.0
        assert math.isclose(float(m[8]), expect, rel_tol=1e-6)  # upper x
        assert m[7] == 0.0 and m[9:13].tolist() == [0.0] * 4
        assert m[0] == m[8]  # bounded x: the pressure is the wall load
        # the fiber shortened from 1 to 0.4: stretch energy with k_s = 3
        m = w.compaction_metrics(1.0, contact_stiffness=0.0, stretch_stiffness=3.0, bending_stiffness=0.0)
        assert math.isclose(float(m[4]), 0.5 * 3.0 * 0.6 ** 2 / 1.0, rel_tol=1e-5)
        assert m[6] == m[4]


def test_compaction_rejects_bad_cells():
    with oracle.OracleWorld(single_fiber(), fx.config(max_step=0.1)) as w:
        with pytest.raises(ValueError):
            w.compact_cell([0.0, 0.0, 0.0], [5.0, 4.0, 4.0], RIGID)  # grows
        with pytest.raises(ValueError):
            w.compact_cell([0.0, 0.0, 0.0], [0.15, 4.0, 4.0], RIGID)  # narrower than 2 r
        with pytest.raises(ValueError):
            w.compact_cell([0.0, 0.0, 0.0], [2.0, 4.0, 4.0], 7)
        assert w.cell_bounds() == ((0.0, 0.0, 0.0), (4.0, 4.0, 4.0))


def test_metrics_closed_forms_on_a_crossing():
    """Two crossing fibers 0.15 apart with radius 0.1: after one measured
    pass the segment penetration is 0.05, so the contact energy is
    k/2 * 0.05^2 for the one segment that stores it; the load sums |corrections|."""
    fibers = [[[1.0, 2.0, 2.0], [3.0, 2.0, 2.0]], [[2.0, 1.0, 2.15], [2.0, 3.0, 2.15]]]
    pk = fx.packed(fibers, 0.1, [4.0] * 3, [True] * 3)
    cfg = fx.config(max_step=0.1, force_full_iterations=1)
    with oracle.OracleWorld(pk, cfg) as w:
        w.run_batch(cfg, 1)
        corr, smax, curv = w.contact_state()
        m = w.compaction_metrics(1.0, contact_stiffness=4.0, stretch_stiffness=0.0, bending_stiffness=0.0)
        # independent float64 sums of the stored state
        assert math.isclose(float(m[3]), sum(0.5 * 4.0 * float(x) ** 2 for x in smax), rel_tol=1e-6)
        load = np.abs(corr.astype(np.float64))
        for k in range(3):  # periodic: pressure from the corrections
            assert math.isclose(float(m[k]), 4.0 * (load[:, k].sum() + load[:, k + 3].sum()) / 16.0,
                                rel_tol=1e-6, abs_tol=1e-12)
        assert m[7:13].tolist() == [0.0] * 6


def test_bending_energy_from_stored_ratio():
    """A bent triplet whose ratio is accepted (tolerance 10) keeps its
    measured ratio; the bending energy is k_b/2 * (ratio - 1)^2 with the
    discrete curvature 2 sin(theta / 2) / dual length."""
    placed = [[[0.8, 1.0, 1.0], [1.0, 1.15, 1.0], [1.2, 1.0, 1.0]]]
    pk = fx.packed(placed, 0.01, [2.0] * 3, [False] * 3, bend_radius=0.25)
    cfg = fx.config(max_step=0.02, curvature_ratio_tolerance=10.0)
    with oracle.OracleWorld(pk, cfg) as w:
        s = w.run_batch(cfg, 5)
        assert s.converged and s.total_iterations == 0
        a = np.array([0.2, 0.15]), np.array([0.2, -0.15])
        l1, l2 = np.linalg.norm(a[0]), np.linalg.norm(a[1])
        theta = math.acos(float(a[0] @ a[1]) / (l1 * l2))
        ratio = 2.0 * math.sin(theta / 2) / (0.5 * (l1 + l2)) * 0.25
        assert ratio > 1.0
        m = w.compaction_metrics(1.0, contact_stiffness=0.0, stretch_stiffness=0.0, bending_stiffness=6.0)
        assert math.isclose(float(m[5]), 0.5 * 6.0 * (ratio - 1.0) ** 2, rel_tol=1e-4)


@pytest.mark.parametrize("builder,gap,tol,step", [(fx.periodic_pair, 0.1, 1e-5, 0.1),
                                                  (fx.si_crossing, 20e-6, 0.01e-6, 1e-6),
                                                  (fx.unwrapped_images, 20e-6, 0.01e-6, 1e-6)])
def test_capture_reference_cases(builder, gap, tol, step):
    # world_tests.rs:934-1026: one candidate each, gap -0.1 / -20 um
    with oracle.OracleWorld(builder(), fx.config(max_step=step)) as w:
        pairs, coords, gaps, angles, overflow = w.capture_contact_arrays(0.0, 16)
        assert not overflow
        assert pairs.shape == (1, 2)
        assert abs(gaps[0] + gap) < tol
        if builder is not fx.periodic_pair:  # perpendicular crossings
            assert math.isclose(float(angles[0]), math.pi / 2, rel_tol=1e-6)
            assert np.allclose(coords[0], [0.5, 0.5], atol=1e-6)
        else:  # parallel
            assert angles[0] == 0.0


def brute_force_capture(positions, counts, radius, lower, upper, periodic, max_gap):
    """All pairs of segments of different fibers, float64, minimum image of
    the midpoint difference, exact segment-segment distance by dense
    sampling refined with the clamped closed form."""
    pts = positions.reshape(-1, 3).astype(np.float64)
    L = np.asarray(upper, np.float64) - np.asarray(lower, np.float64)
    segs, fib = [], []
    start = 0
    for f, n in enumerate(counts):
        for i in range(n - 1):
            segs.append((start + i, start + i + 1))
            fib.append(f)
        start += n
    out = {}
    for i in range(len(segs)):
        for j in range(i + 1, len(segs)):
            if fib[i] == fib[j]:
                continue
            p1, q1 = pts[segs[i][0]], pts[segs[i][1]]
            p2, q2 = pts[segs[j][0]], pts[segs[j][1]]
            mid = 0.5 * (p2 + q2 - p1 - q1)
            for k in range(3):
                if periodic[k]:
                    sh = np.round(mid[k] / L[k]) * L[k]
                    p2 = p2.copy()
                    q2 = q2.copy()
                    p2[k] -= sh
                    q2[k] -= sh
            s = np.linspace(0, 1, 201)
            a = p1[None, :] + s[:, None] * (q1 - p1)[None, :]
            b = p2[None, :] + s[:, None] * (q2 - p2)[None, :]
            d = np.linalg.norm(a[:, None, :] - b[None, :, :], axis=2)
            gap = d.min() - radius[fib[i]] - radius[fib[j]]
            out[(i, j)] = gap
    return out


def test_capture_matches_brute_force_on_a_random_mat():
    rng = np.random.default_rng(7)
    fibers = []
    for _ in range(24):
        c = rng.uniform(0.0, 1.0, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        fibers.append([c + (t - 1.5) * 0.12 * d for t in range(4)])
    radius = np.full(len(fibers), 0.03)
    pk = fx.packed(fibers, radius, [1.0, 1.0, 1.0], [True, True, False])
    max_gap = 0.02
    with oracle.OracleWorld(pk, fx.config(max_step=0.01)) as w:
        pairs, coords, gaps, angles, overflow = w.capture_contact_arrays(max_gap, 4096)
        assert not overflow
        # sorted by (first, second)
        keys = pairs[:, 0].astype(np.int64) * 100000 + pairs[:, 1]
        assert np.all(np.diff(keys) > 0)
        lo, hi = w.cell_bounds()
        bf = brute_force_capture(w.download_positions(), [4] * len(fibers), radius, lo, hi, [True, True, False],
                                 max_gap)
    got = {(int(a), int(b)): float(g) for (a, b), g in zip(pairs, gaps)}
    # the sampled distance overestimates by at most (0.36 / 200)^2 / (8 r)-ish;
    # pairs near the threshold are not decided by the brute force
    margin = 2e-3
    expect = {k for k, g in bf.items() if g <= max_gap - margin}
    undecided = {k for k, g in bf.items() if abs(g - max_gap) <= margin}
    assert expect <= set(got), sorted(expect - set(got))
    assert set(got) <= expect | undecided
    assert len(expect) > 10
    for k in got:
        assert got[k] <= bf[k] + 1e-6 and bf[k] - got[k] < margin


def test_capture_overflow_keeps_capacity():
    with oracle.OracleWorld(fx.crossed(), fx.config(max_step=0.1)) as w:
        pairs, *_, overflow = w.capture_contact_arrays(10.0, 1)
        assert pairs.shape == (1, 2) and not overflow  # one pair exists
    fibers = [[[0.5 + 0.25 * i, 1.0, 1.0], [0.5 + 0.25 * i, 3.0, 1.0]] for i in range(4)]
    with oracle.OracleWorld(fx.packed(fibers, 0.1, [4.0] * 3, [False] * 3), fx.config(max_step=0.1)) as w:
        pairs, *_, overflow = w.capture_contact_arrays(1.0, 2)
        assert pairs.shape == (2, 2) and overflow
        assert pairs.tolist() == [[0, 1], [0, 2]]  # the sorted prefix
