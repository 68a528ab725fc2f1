"""Junction capture's host policy (tangle_amd/junctions.py) against the
reference's unit test (tangle_generate/src/junctions.rs:283-378) and the
policy's rules: angle window, material pairs, deterministic acceptance,
per-pair count and anchor spacing; policy validation as the Python binding
validates it (tangle_python/src/junctions.rs:95-140)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import junctions as jn


class Fiber:
    def __init__(self, material):
        self.material = tg.Material(material, 0.2)


def capture(rows, overflow=False):
    pairs = np.array([[r[0], r[1]] for r in rows], dtype=np.uint32).reshape(-1, 2)
    coords = np.array([[r[2], r[3]] for r in rows], dtype=np.float32).reshape(-1, 2)
    gaps = np.array([r[4] for r in rows], dtype=np.float32)
    angles = np.array([r[5] for r in rows], dtype=np.float32)
    return pairs, coords, gaps, angles, overflow


def two_crossing():
    # junctions.rs:290-305: two 1-long fibers of one segment each
    intrinsic = np.array([[0.5, 1.0, 1.0], [1.5, 1.0, 1.0], [1.0, 0.5, 1.0], [1.0, 1.5, 1.0]], dtype=np.float32)
    return [Fiber("fiber"), Fiber("fiber")], [2, 2], intrinsic


def test_promotes_a_contact_once_with_material_anchors():
    fibers, counts, intrinsic = two_crossing()
    policy = tg.JunctionPolicy("crossing", "welded")
    store = []
    contact = (0, 1, 0.5, 0.5, 0.0, math.pi / 2)
    first = jn.form_junctions(store, fibers, counts, intrinsic, capture([contact]), policy)
    assert first == {"policy_name": "crossing", "candidates": 1, "created": 1, "rejected": 0}
    assert store[0].anchors == ((1, 0.5), (2, 0.5)) and store[0].id == 1 and store[0].law == "welded"
    second = jn.form_junctions(store, fibers, counts, intrinsic, capture([contact]), policy)
    assert second["created"] == 0 and len(store) == 1


def test_policy_rules():
    fibers, counts, intrinsic = two_crossing()
    contact = (0, 1, 0.5, 0.5, 0.0, 0.3)
    store = []
    narrow = tg.JunctionPolicy(min_crossing_angle=0.5)
    assert jn.form_junctions(store, fibers, counts, intrinsic, capture([contact]), narrow)["created"] == 0
    other = tg.JunctionPolicy(material_pairs=[("glass", "fiber")])
    assert jn.form_junctions(store, fibers, counts, intrinsic, capture([contact]), other)["created"] == 0
    never = tg.JunctionPolicy(probability=0.0)
    assert jn.form_junctions(store, fibers, counts, intrinsic, capture([contact]), never)["created"] == 0
    # two contacts of one fiber pair: max_per_fiber_pair decides
    two = [(0, 1, 0.2, 0.5, 0.0, 1.0), (0, 1, 0.8, 0.5, -1e-3, 1.0)]
    assert jn.form_junctions(store, fibers, counts, intrinsic, capture(two), tg.JunctionPolicy())["created"] == 1
    assert store[-1].anchors[0][1] == pytest.approx(0.8, abs=1e-7)  # the deeper contact first
    store2 = []
    report = jn.form_junctions(store2, fibers, counts, intrinsic, capture(two),
                               tg.JunctionPolicy(max_per_fiber_pair=2))
    assert report["created"] == 2 and [j.id for j in store2] == [1, 2]
    store3 = []
    report = jn.form_junctions(store3, fibers, counts, intrinsic, capture(two),
                               tg.JunctionPolicy(max_per_fiber_pair=2, min_anchor_separation=0.7))
    assert report["created"] == 1
    with pytest.raises(OverflowError):
        jn.form_junctions([], fibers, counts, intrinsic, capture([contact], overflow=True), tg.JunctionPolicy())


def test_anchor_arc_length_counts_earlier_segments():
    intrinsic = np.array([[0, 0, 0], [1, 0, 0], [1, 2, 0], [5, 5, 5], [6, 5, 5]], dtype=np.float32)
    assert jn.rest_arc_length(intrinsic, 0, 1, 0.25) == 1.0 + 0.25 * 2.0
    assert jn.rest_arc_length(intrinsic, 3, 3, 1.0) == 1.0


def test_deterministic_unit_is_a_53_bit_fraction():
    values = [jn.deterministic_unit(s, ((1, 0.5), (2, 0.25 * k))) for s in range(5) for k in range(5)]
    assert all(0.0 <= v < 1.0 for v in values)
    assert len(set(values)) == len(values)
    assert jn.deterministic_unit(7, ((3, 1.5), (9, 0.0))) == jn.deterministic_unit(7, ((3, 1.5), (9, 0.0)))
    # a fraction of 53 random bits: the mean of many draws is near 1/2
    draws = [jn.deterministic_unit(s, ((1, 0.0), (2, 0.0))) for s in range(4000)]
    assert abs(sum(draws) / len(draws) - 0.5) < 0.02


def test_policy_validation_and_recording():
    with pytest.raises(ValueError):
        tg.JunctionPolicy(name=" ")
    with pytest.raises(ValueError):
        tg.JunctionPolicy(max_crossing_angle=2.0)
    with pytest.raises(ValueError):
        tg.JunctionPolicy(min_crossing_angle=1.0, max_crossing_angle=0.5)
    with pytest.raises(ValueError):
        tg.JunctionPolicy(probability=1.5)
    with pytest.raises(ValueError):
        tg.JunctionPolicy(candidate_capacity=0)
    cell = tg.Cell([1e-3] * 3, periodic="xyz")
    r = tg.Recipe(cell)
    r.insert(tg.generate_fiber_population(cell, tg.FiberPopulation(material=tg.Material("f", 10e-6), count=2,
                                                                   length=2e-4, segments_per_fiber=4, seed=1)))
    r.capture_junctions(tg.JunctionPolicy("bonds"))
    r.relax_and_capture(iterations=64, capture_every=16, policy=tg.JunctionPolicy())
    assert r.operations()[-2:] == ['capture junctions using policy "bonds"',
                                   "relax for 64 iterations and capture junctions every 16"]
    with pytest.raises(ValueError):
        r.relax_and_capture(iterations=8, capture_every=16, policy=tg.JunctionPolicy())
