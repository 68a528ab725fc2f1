"""Shape edge cases of the HIP path against the oracle, through the C ABI:
a fiber too long for the curvature sweep's LDS staging (the one-lane walk
in global memory takes it), one too long for the sequential 8-lane sweep
(the screened 8-lane sweep takes it), worlds of one-segment fibers (no internal
vertices: no curvature work at all), and a world whose fibers have every
length from one to sixty-four segments."""

from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def _helix(segments, radius, pitch, turn_segments, centre, phase=0.0):
    points = []
    for i in range(segments + 1):
        a = phase + 2.0 * math.pi * i / turn_segments
        points.append([centre[0] + radius * math.cos(a), centre[1] + radius * math.sin(a),
                       centre[2] + pitch * i / turn_segments])
    return points


def _run_both(packed, cfg, batches, iterations):
    with _native.World(packed, cfg) as gpu, oracle.OracleWorld(packed, cfg) as ref:
        lanes = gpu.tuning().curvature_lanes
        for b in range(batches):
            print(f"[case] batch {b} of {iterations} iterations", flush=True)
            sg = gpu.run_batch(cfg, iterations)
            so = ref.run_batch(cfg, iterations)
            assert sg.total_iterations == so.total_iterations
            assert sg.converged == so.converged
            assert np.float32(sg.max_penetration) == np.float32(so.max_penetration)
            assert np.float32(sg.max_curvature_ratio) == np.float32(so.max_curvature_ratio)
            assert np.float32(sg.max_displacement) == np.float32(so.max_displacement)
            assert sg.cell_list_overflow == so.cell_list_overflow
        return lanes, gpu.download_positions(), ref.download_positions()


def test_fiber_longer_than_lds_takes_the_global_walk(hip_device):
    """A 1,800-segment coil (1,801 vertices, beyond the 1,536 an LDS-staged
    sweep holds) whose turns overlap and whose bend radius is below the
    limit, with two short coils beside it: the world chooses the one-lane
    walk, and positions and status match the oracle bit for bit."""
    fibers = [_helix(1800, 0.1, 0.008, 40, (0.5, 0.5, 0.1)),
              _helix(30, 0.05, 0.02, 12, (0.2, 0.8, 0.5), 0.3),
              _helix(12, 0.05, 0.02, 12, (0.8, 0.2, 0.5), 1.1)]
    packed = fx.packed(fibers, 0.005, [1.0] * 3, [True] * 3, bend_radius=0.2)
    cfg = fx.config(force_full_iterations=1, max_step=0.002)
    lanes, pg, po = _run_both(packed, cfg, 2, 15)
    assert lanes == 0
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


def test_long_coil_in_the_sequential_sweep_matches_bitwise(hip_device):
    """A 200-segment coil (201 vertices, three times the felt's longest
    fiber) with two short coils, in the 8-lane sequential sweep: positions
    and status match the oracle bit for bit."""
    fibers = [_helix(200, 0.1, 0.01, 40, (0.5, 0.5, 0.3)),
              _helix(30, 0.05, 0.02, 12, (0.2, 0.8, 0.5), 0.3),
              _helix(12, 0.05, 0.02, 12, (0.8, 0.2, 0.5), 1.1)]
    packed = fx.packed(fibers, 0.005, [1.0] * 3, [True] * 3, bend_radius=0.2)
    cfg = fx.config(force_full_iterations=1, max_step=0.002)
    lanes, pg, po = _run_both(packed, cfg, 2, 15)
    assert lanes == 8
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


def test_fiber_too_long_for_the_sequential_sweep_takes_a_wider_group(hip_device):
    """A 420-segment coil (421 vertices: more than eight fibers' records of
    the sequential sweep fit in 64 KiB of LDS) in a periodic box: the world
    picks 16 lanes per fiber (the screened sweep), and positions and status
    match the oracle bit for bit."""
    fibers = [_helix(420, 0.1, 0.005, 60, (0.5, 0.5, 0.25)),
              _helix(30, 0.05, 0.02, 12, (0.2, 0.8, 0.5), 0.3)]
    packed = fx.packed(fibers, 0.004, [1.0] * 3, [True] * 3, bend_radius=0.2)
    cfg = fx.config(force_full_iterations=1, max_step=0.002)
    lanes, pg, po = _run_both(packed, cfg, 2, 15)
    assert lanes == 16
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


def test_one_segment_fibers_match_bitwise(hip_device):
    """300 straight one-segment rods at random in a periodic box (many
    overlapping): only contact and stretch act."""
    rng = np.random.default_rng(11)
    fibers = []
    for _ in range(300):
        c = rng.uniform(0.0, 1.0, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        fibers.append([list(c - 0.03 * d), list(c + 0.03 * d)])
    packed = fx.packed(fibers, 0.01, [1.0] * 3, [True] * 3, bend_radius=0.05)
    cfg = fx.config(force_full_iterations=1, max_step=0.005)
    _, pg, po = _run_both(packed, cfg, 3, 10)
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


def test_every_fiber_length_from_one_to_sixty_four(hip_device):
    """64 coils of 1 to 64 segments, interleaved in a closed box (walls
    live), bend-limited: ragged fibers share curvature wavefronts."""
    rng = np.random.default_rng(5)
    fibers = []
    for n in rng.permutation(np.arange(1, 65)):
        c = rng.uniform(0.25, 0.75, 3)
        fibers.append(_helix(int(n), 0.04, 0.03, 10, (c[0], c[1], c[2] - 0.0015 * n), float(rng.uniform(0, 6))))
    packed = fx.packed(fibers, 0.006, [1.0] * 3, [False] * 3, bend_radius=0.06)
    cfg = fx.config(force_full_iterations=1, max_step=0.003)
    lanes, pg, po = _run_both(packed, cfg, 2, 20)
    assert lanes == 8
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))
