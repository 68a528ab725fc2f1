"""Junction capture: promoting captured fiber-fiber contacts to persistent
junction records (tangle_generate/src/junctions.rs:8-283; the Python policy
of tangle_python/src/junctions.rs).

The device captures the candidate pairs (`World.capture_contact_arrays`,
sorted as the reference sorts them); this host step filters them by
crossing angle, material pair, a seeded deterministic acceptance draw, the
per-fiber-pair count and the spacing from existing junctions, and appends
the accepted ones. Junctions are recipe results: they do not constrain the
relaxation.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace as _replace

import numpy as np

MASK64 = (1 << 64) - 1
HALF_PI_F32 = float(np.float32(math.pi / 2))


@dataclass
class JunctionPolicy:
    """Deterministic rules for promoting current contacts to junctions
    (JunctionPolicy; defaults of JunctionCapturePolicy::touching)."""

    name: str = "touching contacts"
    law_name: str = "bond"
    parameter_set: int = 0
    max_surface_gap: float = 1.0e-4
    min_crossing_angle: float = 0.0
    max_crossing_angle: float = HALF_PI_F32
    probability: float = 1.0
    seed: int = 0
    material_pairs: list = field(default_factory=list)
    max_per_fiber_pair: int = 1
    min_anchor_separation: float = 1.0e-6
    candidate_capacity: int = 65_536

    def __post_init__(self):
        self.validate()

    def replace(self, **changes) -> "JunctionPolicy":
        return _replace(self, **changes)

    def validate(self) -> None:
        if not str(self.name).strip() or not str(self.law_name).strip():
            raise ValueError("junction policy and law names must not be empty")
        if not (math.isfinite(self.max_surface_gap) and self.max_surface_gap >= 0.0):
            raise ValueError("max_surface_gap must be nonnegative and finite")
        if (not math.isfinite(self.min_crossing_angle) or not math.isfinite(self.max_crossing_angle)
                or self.min_crossing_angle < 0.0 or np.float32(self.max_crossing_angle) > np.float32(math.pi / 2)
                or self.min_crossing_angle > self.max_crossing_angle):
            raise ValueError("crossing angles must be ordered within [0, pi/2] radians")
        if not (math.isfinite(self.probability) and 0.0 <= self.probability <= 1.0):
            raise ValueError("junction probability must be in [0, 1]")
        if int(self.max_per_fiber_pair) <= 0 or int(self.candidate_capacity) <= 0:
            raise ValueError("junction count and candidate capacity must be positive")
        if not (math.isfinite(self.min_anchor_separation) and self.min_anchor_separation >= 0.0):
            raise ValueError("min_anchor_separation must be nonnegative and finite")
        if not 0 <= int(self.seed) <= MASK64:
            raise ValueError("seed must be an unsigned 64-bit integer")
        self.material_pairs = [(str(a), str(b)) for a, b in self.material_pairs]


@dataclass(frozen=True)
class Junction:
    """One persistent junction: its id, law name, parameter set and the two
    anchors (fiber id, rest arc length), lower fiber id first."""

    id: int
    law: str
    parameter_set: int
    anchors: tuple


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & MASK64


def _f64_bits(x: float) -> int:
    return int(np.float64(x).view(np.uint64))


def deterministic_unit(seed: int, anchors) -> float:
    """A SplitMix64-style finalisation of the seed and both anchors, mapped to
    [0, 1) with 53 bits (deterministic_unit, junctions.rs:262-277)."""
    (f0, s0), (f1, s1) = anchors
    v = (int(seed) ^ ((f0 * 0x9E3779B97F4A7C15) & MASK64) ^ ((f1 * 0xBF58476D1CE4E5B9) & MASK64)
         ^ _rotl(_f64_bits(s0), 17) ^ _rotl(_f64_bits(s1), 41)) & MASK64
    v ^= v >> 30
    v = (v * 0xBF58476D1CE4E5B9) & MASK64
    v ^= v >> 27
    v = (v * 0x94D049BB133111EB) & MASK64
    v ^= v >> 31
    return float(v >> 11) * (1.0 / float(1 << 53))


def _total_key(x) -> int:
    i = int(np.float32(x).view(np.int32))
    return i ^ ((i >> 31) & 0x7FFFFFFF)


def rest_arc_length(intrinsic_f32: np.ndarray, fiber_start: int, first_vertex: int, coordinate) -> float:
    """Rest arc length of a point of a segment, from the f32 intrinsic
    positions: whole segments before it, then the coordinate's share of its
    own (anchor_for_segment, junctions.rs:197-222)."""
    def dist(a, b):
        d = intrinsic_f32[b] - intrinsic_f32[a]
        return float(np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))

    total = 0.0
    for v in range(fiber_start, first_vertex):
        total += dist(v, v + 1)
    total += float(np.float32(coordinate)) * dist(first_vertex, first_vertex + 1)
    return total


def form_junctions(junctions: list, fibers, counts, intrinsic_f32: np.ndarray, capture, policy: JunctionPolicy):
    """Appends the accepted junctions of one capture to `junctions` and
    returns the report (form_junctions_from_capture, junctions.rs:100-162).
    `capture` = (pairs, coordinates, gaps, angles, overflow) as
    World.capture_contact_arrays returns it; fiber i has id i + 1."""
    pairs, coords, gaps, angles, overflow = capture
    if overflow:
        raise OverflowError(f"junction policy '{policy.name}' exceeded its candidate or broad-phase capacity")
    counts = np.asarray(counts, dtype=np.int64)
    vstart = np.concatenate([[0], np.cumsum(counts)[:-1]])
    sstart = np.concatenate([[0], np.cumsum(counts - 1)[:-1]])
    n = len(gaps)
    order = sorted(range(n), key=lambda i: (_total_key(gaps[i]), int(pairs[i][0]), int(pairs[i][1]),
                                            _total_key(coords[i][0]), _total_key(coords[i][1])))
    lo_angle, hi_angle = np.float32(policy.min_crossing_angle), np.float32(policy.max_crossing_angle)
    probability = float(np.float32(policy.probability))
    tolerance = max(float(policy.min_anchor_separation), 1.0e-9)
    created = 0
    for i in order:
        angle = np.float32(angles[i])
        if angle < lo_angle or angle > hi_angle:
            continue
        anchors = []
        for side in range(2):
            seg = int(pairs[i][side])
            fiber = int(np.searchsorted(sstart, seg, side="right") - 1)
            first_vertex = int(vstart[fiber] + (seg - sstart[fiber]))
            anchors.append((fiber + 1, rest_arc_length(intrinsic_f32, int(vstart[fiber]), first_vertex,
                                                       coords[i][side])))
        if policy.material_pairs:
            a = fibers[anchors[0][0] - 1].material.name
            b = fibers[anchors[1][0] - 1].material.name
            if not any((x == a and y == b) or (x == b and y == a) for x, y in policy.material_pairs):
                continue
        if anchors[1][0] < anchors[0][0]:
            anchors.reverse()
        anchors = tuple(anchors)
        if deterministic_unit(policy.seed, anchors) > probability:
            continue
        same_pair = [j for j in junctions if {j.anchors[0][0], j.anchors[1][0]} == {anchors[0][0], anchors[1][0]}
                     and anchors[0][0] != anchors[1][0]]
        if len(same_pair) >= policy.max_per_fiber_pair:
            continue
        near = False
        for j in junctions:
            ordered = j.anchors if j.anchors[0][0] <= j.anchors[1][0] else (j.anchors[1], j.anchors[0])
            if (ordered[0][0] == anchors[0][0] and ordered[1][0] == anchors[1][0]
                    and abs(ordered[0][1] - anchors[0][1]) <= tolerance
                    and abs(ordered[1][1] - anchors[1][1]) <= tolerance):
                near = True
                break
        if near:
            continue
        next_id = max((j.id for j in junctions), default=0) + 1
        junctions.append(Junction(next_id, policy.law_name, int(policy.parameter_set), anchors))
        created += 1
    return {"policy_name": policy.name, "candidates": n, "created": created, "rejected": n - created}
