"""Host side of the formation recipe operations: layer-centre targets, needle
footprints and the selection of needled vertices.

These run between device batches, on the host, as the reference's recipe
controller does (crates/tangle_generate/src/recipe.rs:857-1205, 1688-1812);
the targets they produce are installed on the resident device world
(tgh_world_set_layer_targets / tgh_world_set_vertex_targets) and applied there
after every iteration.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

MASK = (1 << 64) - 1
NO_LAYER = 0xFFFFFFFF


def splitmix64(value: int) -> int:
    """One SplitMix64 output from state `value` (recipe.rs:1814-1819)."""
    value = (value + 0x9E3779B97F4A7C15) & MASK
    value = ((value ^ (value >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    value = ((value ^ (value >> 27)) * 0x94D049BB133111EB) & MASK
    return value ^ (value >> 31)


def random_footprint_center(seed: int, layer: int, origin, extent) -> tuple[np.float32, np.float32]:
    """A reproducible footprint centre in `origin + [0, extent)` over the two
    plane coordinates (recipe.rs:230-246): 24 random bits per coordinate,
    f32 arithmetic."""
    def unit(bits):
        return np.float32(float(bits >> 40) * (1.0 / float(1 << 24)))

    first = splitmix64((seed ^ (2 * layer)) & MASK)
    second = splitmix64((seed ^ (2 * layer + 1)) & MASK)
    o = np.asarray(origin, dtype=np.float32)
    e = np.asarray(extent, dtype=np.float32)
    return (np.float32(o[0] + np.float32(e[0] * unit(first))), np.float32(o[1] + np.float32(e[1] * unit(second))))


def layer_targets(centerlines, layers, axis: int) -> list[np.float32]:
    """Initial centre of every formation layer along `axis`: the mean over the
    layer's fibers of each fiber's mean vertex coordinate, in f64, rounded to
    f32 (crates/tangle_generate/src/layered.rs:178-213)."""
    labelled = [layer for layer in layers if layer is not None]
    if not labelled:
        raise ValueError("layered formation requires formation-layer labels")
    count = max(labelled) + 1
    sums = [0.0] * count
    counts = [0] * count
    for line, layer in zip(centerlines, layers):
        if layer is None:
            continue
        sums[layer] += _ordered_sum([float(p[axis]) for p in line]) / len(line)
        counts[layer] += 1
    out = []
    for s, c in zip(sums, counts):
        if c == 0:
            raise ValueError("formation layers must be contiguous and populated")
        out.append(np.float32(s / c))
    return out


def _ordered_sum(values) -> float:
    """Left-to-right f64 sum, as Rust's Iterator::sum."""
    total = 0.0
    for v in values:
        total += v
    return total


def round_half_away(x: np.ndarray) -> np.ndarray:
    """f32::round: halves away from zero (numpy rounds halves to even)."""
    t = np.trunc(x)
    frac = x - t
    return (t + np.where(np.abs(frac) >= np.float32(0.5), np.sign(x), np.float32(0.0))).astype(x.dtype)


@dataclass(frozen=True)
class CircularFootprint:
    """Pull fibers whose centreline crosses a circle in the layer plane
    (crates/tangle_python/src/recipe.rs:64-121). `CircularFootprint(center,
    diameter=...)` places the circle; `CircularFootprint.random(diameter=...,
    seed=...)` picks a reproducible centre per layer over the cell footprint."""

    center: tuple | None
    diameter: float
    seed: int | None = None

    def __init__(self, center, *, diameter: float, seed: int | None = None):
        if center is not None:
            center = tuple(float(c) for c in center)
            if len(center) != 2 or not all(math.isfinite(c) for c in center):
                raise ValueError("footprint center must be two finite coordinates")
        if not (math.isfinite(diameter) and diameter > 0.0):
            raise ValueError("diameter must be positive and finite")
        object.__setattr__(self, "center", center)
        object.__setattr__(self, "diameter", float(diameter))
        object.__setattr__(self, "seed", None if seed is None else int(seed))

    @classmethod
    def random(cls, *, diameter: float, seed: int) -> "CircularFootprint":
        return cls(None, diameter=diameter, seed=seed)


@dataclass(frozen=True)
class RandomFiberFraction:
    """Pull one internal vertex of a seeded random `fraction` of the layer's
    fibers (crates/tangle_python/src/recipe.rs:123-146)."""

    fraction: float
    seed: int = 0

    def __post_init__(self):
        if not (math.isfinite(self.fraction) and 0.0 < self.fraction <= 1.0):
            raise ValueError("fraction must lie in (0, 1]")


@dataclass(frozen=True)
class Needling:
    """NeedlingConfig (recipe.rs:268-288) with the footprint resolved: a
    (center, diameter) circle or a RandomFiberFraction."""

    layer: int
    selection: object
    min_fiber_diameter: float | None
    depth: float
    stiffness: float
    max_translation: float
    max_translation_over_diameter: float


def select_needle_vertices(fibers, positions: np.ndarray, starts, active_through, layer_axis: int,
                           periodic, lower, upper, needling: Needling) -> list[int]:
    """selected_needling_vertices_from_mask (recipe.rs:1710-1812) for
    non-adaptive layouts, where a fiber's vertices and segments are active
    exactly when the fiber is. `fibers` are the recipe's fibers in packed
    order (fiber i has id i + 1, as the Python model numbers them), `starts`
    their first vertex, `positions` the resident f32 positions (V, 3),
    `lower`/`upper` the packed f32 cell bounds (packed.rs:306-311)."""
    plane = [k for k in range(3) if k != layer_axis]
    lower = np.asarray(lower, dtype=np.float32)
    upper = np.asarray(upper, dtype=np.float32)
    out: list[int] = []
    for index, fiber in enumerate(fibers):
        if fiber.formation_layer != needling.layer:
            continue
        if active_through is not None and fiber.step > active_through:
            continue
        if needling.min_fiber_diameter is not None and \
                2.0 * fiber.radius < float(np.float32(needling.min_fiber_diameter)):
            continue
        start = int(starts[index])
        count = fiber.placed.shape[0]
        sel = needling.selection
        if isinstance(sel, RandomFiberFraction):
            if count <= 2:
                continue
            selection = splitmix64(np.uint64(int(sel.seed)).item() ^ (index + 1))
            unit = float(selection >> 11) * (1.0 / float(1 << 53))
            if unit >= float(np.float32(sel.fraction)):
                continue
            local = 1 + splitmix64(selection) % (count - 2)
            out.append(start + local)
            continue
        center, diameter = sel
        pts = positions[start:start + count]
        first = pts[:-1]
        second = pts[1:]
        delta = np.empty((count - 1, 2), dtype=np.float32)
        direction = np.empty((count - 1, 2), dtype=np.float32)
        for j, axis in enumerate(plane):
            delta[:, j] = first[:, axis] - center[j]
            direction[:, j] = second[:, axis] - first[:, axis]
            if periodic[axis]:
                width = upper[axis] - lower[axis]
                delta[:, j] -= round_half_away(delta[:, j] / width) * width
                direction[:, j] -= round_half_away(direction[:, j] / width) * width
        length2 = direction[:, 0] * direction[:, 0] + direction[:, 1] * direction[:, 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = -(delta[:, 0] * direction[:, 0] + delta[:, 1] * direction[:, 1]) / length2
        coordinate = np.where(length2 > np.float32(1e-20), np.clip(t, np.float32(0.0), np.float32(1.0)),
                              np.float32(0.0)).astype(np.float32)
        dx = delta[:, 0] + coordinate * direction[:, 0]
        dy = delta[:, 1] + coordinate * direction[:, 1]
        radius2 = dx * dx + dy * dy
        limit = np.float32(np.float32(0.25) * diameter) * diameter
        inside = np.nonzero(radius2 <= limit)[0]
        if inside.size == 0:
            continue
        best = inside[np.argmin(radius2[inside])]  # the first of equal minima
        a = np.float32(start + best)
        material = int(round_half_away(np.array([a + coordinate[best] * np.float32(1.0)], dtype=np.float32))[0])
        out.append(min(max(material, start + 1), start + count - 2))
    return out
