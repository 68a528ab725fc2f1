"""Host side of the formation operations (CPU): layer-centre targets, the
reproducible needle footprint and the needled-vertex selection, against the
reference's own tests (crates/tangle_generate/src/layered.rs:221-240,
crates/tangle_generate/src/recipe_tests.rs:182-336), plus the Recipe
vocabulary that records them (crates/tangle_python/src/recipe.rs:197-790)."""

from __future__ import annotations

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import formation
from tangle_amd import _Fiber


def test_layer_targets_known_answer():
    # layered.rs:221-240
    lines = [np.array([[0.4, 0.5, 0.2], [0.6, 0.5, 0.2]]), np.array([[0.4, 0.5, 0.8], [0.6, 0.5, 0.8]])]
    assert formation.layer_targets(lines, [0, 1], 2) == [np.float32(0.2), np.float32(0.8)]
    # the mean of fiber means, unlabelled fibers ignored
    lines.append(np.array([[0.0, 0.0, 0.3], [0.0, 0.0, 0.5]]))
    lines.append(np.array([[0.0, 0.0, 0.9], [0.0, 0.0, 0.1]]))
    got = formation.layer_targets(lines, [0, 1, 0, None], 2)
    assert got == [np.float32((0.2 + 0.4) / 2), np.float32(0.8)]
    with pytest.raises(ValueError):
        formation.layer_targets(lines, [None] * 4, 2)
    with pytest.raises(ValueError):
        formation.layer_targets(lines[:2], [0, 2], 2)  # layer 1 empty


def test_random_footprint_center_golden_and_bounds():
    # recipe_tests.rs:312-336, pinned against the felt example
    c = formation.random_footprint_center(20_260_940, 2, [0.0, 0.0], [1.0e-3, 1.0e-3])
    assert (c[0], c[1]) == (np.float32(0.000_797_881_1), np.float32(0.000_692_911_7))
    for layer in (0, 1, 7, 0xFFFFFFFF):
        a = formation.random_footprint_center(11, layer, [-1.0, 2.0], [0.5, 0.25])
        assert a == formation.random_footprint_center(11, layer, [-1.0, 2.0], [0.5, 0.25])
        assert -1.0 <= a[0] < -0.5 and 2.0 <= a[1] < 2.25
    assert formation.random_footprint_center(11, 0, [0, 0], [1, 1]) != \
        formation.random_footprint_center(11, 1, [0, 0], [1, 1])


def test_splitmix64_published_value():
    # Vigna's splitmix64 from state 0 yields 0xe220a8397b1dcdaf first
    assert formation.splitmix64(0) == 0xE220A8397B1DCDAF


def _fibers(spec):
    out = []
    for layer, step, radius, pts in spec:
        p = np.asarray(pts, dtype=np.float64)
        out.append(_Fiber(tg.Material("f", 2 * radius), radius, p, p.copy(), layer, step))
    return out


def _select(fibers, needling, active_through, periodic=(False, False, False), cell=1.0):
    positions = np.concatenate([f.placed for f in fibers]).astype(np.float32)
    starts = np.concatenate([[0], np.cumsum([f.placed.shape[0] for f in fibers])[:-1]])
    return formation.select_needle_vertices(fibers, positions, starts, active_through, 2, periodic,
                                            [0.0] * 3, [cell] * 3, needling)


def test_random_fraction_selection_is_reproducible_internal_and_layer_scoped():
    # recipe_tests.rs:182-261
    pts = [[0.2 + 0.1 * i, 0.2, 0.2] for i in range(4)]
    fibers = _fibers([(0, 0, 0.01, pts), (1, 1, 0.01, pts), (1, 1, 0.01, pts)])
    needling = formation.Needling(1, formation.RandomFiberFraction(1.0, seed=42), None, 0.1, 1.0, 0.1, 0.25)
    first = _select(fibers, needling, 1)
    assert first == _select(fibers, needling, 1)
    assert len(first) == 2
    for v in first:
        f = v // 4
        assert fibers[f].formation_layer == 1
        assert 4 * f < v < 4 * f + 3
    # fibers not yet inserted are not pulled
    assert _select(fibers, needling, 0) == []


def test_random_fraction_matches_its_probability():
    pts = [[0.1 * i, 0.5, 0.5] for i in range(6)]
    fibers = _fibers([(0, 0, 0.01, pts)] * 4000)
    needling = formation.Needling(0, formation.RandomFiberFraction(0.3, seed=7), None, 0.1, 1.0, 0.1, 0.25)
    chosen = _select(fibers, needling, 0)
    assert 0.27 < len(chosen) / 4000 < 0.33
    local = np.array(chosen) % 6
    assert local.min() >= 1 and local.max() <= 4 and len(set(local.tolist())) == 4


def test_circular_selection_known_answer():
    # recipe_tests.rs:263-310: only the thick layer-1 fiber through the
    # circle qualifies, and its vertex nearest the centre (6) is pulled
    spec = []
    for layer, y, r in ((1, 0.5, 0.01), (1, 0.5, 0.02), (0, 0.5, 0.02), (1, 0.8, 0.02)):
        spec.append((layer, layer, r, [[0.3, y, 0.5], [0.4, y, 0.5], [0.5, y, 0.5], [0.6, y, 0.5]]))
    fibers = _fibers(spec)
    circle = ((np.float32(0.5), np.float32(0.5)), np.float32(0.1))
    needling = formation.Needling(1, circle, 0.03, 0.1, 1.0, 0.1, 0.25)
    assert _select(fibers, needling, 1) == [6]
    # without the diameter bound the thin fiber joins (vertex 2)
    needling = formation.Needling(1, circle, None, 0.1, 1.0, 0.1, 0.25)
    assert _select(fibers, needling, 1) == [2, 6]


def test_circular_selection_wraps_periodic_axes_and_clamps_to_internal():
    # a fiber across the periodic x face: the circle at x = 0.02 finds the
    # segment whose minimum image passes through it; a circle at a fiber end
    # pulls the first internal vertex
    pts = [[0.9 + 0.04 * i, 0.5, 0.5] for i in range(6)]  # 0.90 .. 1.10
    fibers = _fibers([(0, 0, 0.01, pts)])
    near = ((np.float32(0.02), np.float32(0.5)), np.float32(0.02))
    needling = formation.Needling(0, near, None, 0.1, 1.0, 0.1, 0.25)
    assert _select(fibers, needling, 0, periodic=(True, True, False)) == [3]
    assert _select(fibers, needling, 0, periodic=(False, False, False)) == []
    end = ((np.float32(0.9), np.float32(0.5)), np.float32(0.02))
    assert _select(fibers, formation.Needling(0, end, None, 0.1, 1.0, 0.1, 0.25), 0) == [1]


def test_round_half_away_from_zero():
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, 0.49999997, -2.4], dtype=np.float32)
    assert formation.round_half_away(x).tolist() == [1.0, 2.0, 3.0, -1.0, -2.0, 0.0, -2.0]


def test_recipe_records_formation_operations():
    cell = tg.Cell([1e-3, 1e-3, 0.5e-3], periodic="xy")
    recipe = tg.Recipe(cell)
    fine = tg.Material("fine", 7e-6, 200e-6)
    for ply in range(2):
        col = tg.FiberCollection.from_centerlines([[[1e-4, 5e-4, 1e-4 + 2e-4 * ply], [9e-4, 5e-4, 1e-4 + 2e-4 * ply]]],
                                                  fine, formation_layer=ply)
        recipe.insert(col, name=f"ply {ply}")
    with recipe.place_layer_above(1, gap=2e-4) as r:
        assert r is recipe
        recipe.settle_targets(tolerance=1e-6, max_iterations=100)
    held = recipe.needle_layer(1, footprint=tg.CircularFootprint.random(diameter=2e-4, seed=3), depth=3e-5)
    assert isinstance(held, tg.HeldTargets)
    recipe.release_needles()
    ops = recipe.operations()
    assert ops[0].startswith("insert 'ply 0'") and "formation step 1" in ops[1]
    assert ops[2] == "place layer 1 above layer 0"
    assert ops[4] == "release layer placement"
    assert ops[5].startswith("needle layer 1 through a 2.000e-04 circle")
    packed = recipe._packed(False)
    assert packed.formation_steps.tolist() == [0, 1] and packed.formation_layers.tolist() == [0, 1]
    with pytest.raises(ValueError):
        recipe.place_layer_above(0, gap=1e-4)
    with pytest.raises(ValueError):
        recipe.needle_layer(1, footprint=tg.RandomFiberFraction(0.5), depth=-1.0)
    with pytest.raises(TypeError):
        recipe.needle_layer(1, footprint="circle", depth=1e-5)
    with pytest.raises(ValueError):
        tg.RandomFiberFraction(1.5)
