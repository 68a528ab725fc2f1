"""TANGLE relaxation on the MI355X: the reference's recipe API over a HIP backend.

This package mirrors the part of TANGLE's Python API that drives the
relaxation hot path (crates/tangle_python/python/tangle/_tangle.pyi):
``Cell``, ``Material``, ``FiberPopulation`` with its orientation classes,
``generate_fiber_population``, ``FiberCollection``, ``Recipe`` with
``insert`` / ``relax_until_converged`` / ``relax_for`` / ``run``,
``RelaxationSettings`` and ``RunResult``. ``RelaxationSettings(backend="hip")``
selects the hand-written CDNA4 kernels of ``libtangle_hip.so``; it is the only
backend this package builds, and there is no silent fallback: other backend
names raise ``ValueError`` and a missing library raises ``RuntimeError``.

Lengths are SI metres, as in the reference.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace as _dc_replace
from typing import Iterable, Sequence

import numpy as np

from . import _native
from .checkpoint import CheckpointError, CheckpointSettings
from .compaction import (AxisWeightsPath, CellLengthsTarget, CellVolumeTarget, CompactionSettings,
                         DirectionalPressureTarget, EqualPressurePath, MeanPressureTarget, MinimumWorkPath,
                         PenaltyEnergyTarget, StressRatioPath, VolumeFractionTarget)
from .formation import CircularFootprint, RandomFiberFraction
from .junctions import JunctionPolicy
from .units import um

__all__ = [
    "Cell",
    "Material",
    "FiberCollection",
    "FiberPopulation",
    "IsotropicOrientation",
    "PlanarOrientation",
    "AlignedOrientation",
    "SchladitzOrientation",
    "LayeredBiaxialOrientation",
    "UniformPosition",
    "LayeredPosition",
    "DensityGradientPosition",
    "generate_fiber_population",
    "RelaxationSettings",
    "RelaxationOverrides",
    "SolvePolicy",
    "Recipe",
    "RecipeError",
    "RunResult",
    "CircularFootprint",
    "RandomFiberFraction",
    "HeldTargets",
    "CheckpointSettings",
    "CheckpointError",
    "relax_batch",
    "CompactionSettings",
    "JunctionPolicy",
    "VolumeFractionTarget",
    "CellVolumeTarget",
    "CellLengthsTarget",
    "MeanPressureTarget",
    "DirectionalPressureTarget",
    "PenaltyEnergyTarget",
    "AxisWeightsPath",
    "EqualPressurePath",
    "StressRatioPath",
    "MinimumWorkPath",
    "um",
]

_AXES = {"x": 0, "y": 1, "z": 2, 0: 0, 1: 1, 2: 2}


def _rust_debug(text: str) -> str:
    """A string as Rust's `{:?}` writes it (plain names: in double quotes)."""
    return '"' + text.replace("\\", "\\\\").replace('"', '\\"') + '"'


def _rust_display(value: float) -> str:
    """An f64 as Rust's `{}` writes it: shortest round trip, never exponent form."""
    return np.format_float_positional(float(value), unique=True, trim="-")


def _axis_set(value) -> list[bool]:
    if isinstance(value, str):
        flags = [False, False, False]
        for letter in value.lower():
            if letter not in "xyz":
                raise ValueError(f"unknown axis {letter!r} in {value!r}")
            flags[_AXES[letter]] = True
        return flags
    if isinstance(value, int):
        flags = [False, False, False]
        flags[_AXES[value]] = True
        return flags
    flags = [bool(v) for v in value]
    if len(flags) != 3:
        raise ValueError("an axis set needs three flags")
    return flags


def _direction(value) -> list[float]:
    if isinstance(value, (str, int)):
        d = [0.0, 0.0, 0.0]
        d[_AXES[value]] = 1.0
        return d
    d = [float(v) for v in value]
    if len(d) != 3 or not all(math.isfinite(v) for v in d) or not any(d):
        raise ValueError("orientation direction must be finite and nonzero")
    return d


class Cell:
    """Orthorhombic simulation box (``lengths`` in metres, ``periodic`` axes)."""

    def __init__(self, lengths: Sequence[float], periodic="", origin: Sequence[float] = (0.0, 0.0, 0.0), *,
                 stack_axis=None):
        self._lengths = [float(v) for v in lengths]
        if len(self._lengths) != 3 or not all(v > 0 and math.isfinite(v) for v in self._lengths):
            raise ValueError("cell lengths must be three positive finite numbers")
        self._periodic = _axis_set(periodic)
        self._origin = [float(v) for v in origin]
        self._stack_axis = 2 if stack_axis is None else _AXES[stack_axis]

    lengths = property(lambda self: list(self._lengths))
    periodic = property(lambda self: list(self._periodic))
    origin = property(lambda self: list(self._origin))
    stack_axis = property(lambda self: self._stack_axis)

    def __repr__(self) -> str:
        axes = "".join("xyz"[k] for k in range(3) if self._periodic[k])
        return f"Cell(lengths={self._lengths}, periodic={axes!r})"


class Material:
    """A fiber type: ``diameter`` in metres and an optional ``min_bend_radius``."""

    def __init__(self, name: str, diameter: float, min_bend_radius: float | None = None, *,
                 thickness: float | None = None):
        if not (diameter > 0 and math.isfinite(diameter)):
            raise ValueError("diameter must be positive and finite")
        if min_bend_radius is not None and not (min_bend_radius > 0 and math.isfinite(min_bend_radius)):
            raise ValueError("min_bend_radius must be positive and finite")
        self._name = str(name)
        self._diameter = float(diameter)
        self._min_bend_radius = None if min_bend_radius is None else float(min_bend_radius)
        self._thickness = None if thickness is None else float(thickness)

    name = property(lambda self: self._name)
    diameter = property(lambda self: self._diameter)
    radius = property(lambda self: 0.5 * self._diameter)
    min_bend_radius = property(lambda self: self._min_bend_radius)
    thickness = property(lambda self: self._thickness)
    is_oval = property(lambda self: self._thickness is not None and self._thickness < self._diameter)

    def __repr__(self) -> str:
        return f'Material(name="{self._name}", diameter={self._diameter!r})'


@dataclass(frozen=True)
class IsotropicOrientation:
    """Uniform directions on the sphere (OrientationDistribution::Isotropic3d)."""


@dataclass(frozen=True)
class PlanarOrientation:
    """Directions within ``max_tilt`` radians of the plane normal to ``normal``."""

    normal: object = "z"
    max_tilt: float = 0.0


@dataclass(frozen=True)
class AlignedOrientation:
    """Directions in a cone of half-angle ``max_angle`` around ``axis``."""

    axis: object = "z"
    max_angle: float = 0.0


@dataclass(frozen=True)
class SchladitzOrientation:
    """Schladitz orientation density β/(4π(1+(β²−1)cos²θ)^{3/2}) about ``axis``.

    ``beta`` follows the reference's fit convention
    (crates/tangle_characterize/src/shape.rs:26-30): β = 1 is isotropic,
    β → 0 aligns fibers with ``axis`` and β > 1 lays them in the plane
    normal to it. The reference only fits β; this sampler is an addition.
    """

    beta: float = 1.0
    axis: object = "z"


@dataclass(frozen=True)
class LayeredBiaxialOrientation:
    """Per-layer biaxial directions in the plane normal to ``normal``
    (OrientationDistribution::LayeredBiaxial, population.rs:599-628): a
    ``primary_fraction`` within ``max_in_plane_deviation`` of the layer's
    seeded reference angle, a ``cross_fraction`` about the angle a quarter
    turn away, the rest uniform; tilted by at most ``max_tilt``. ``normal``
    None is the stack axis. Needs ``LayeredPosition`` (generators.rs:96-125)."""

    normal: object = None
    primary_fraction: float = 0.4
    cross_fraction: float = 0.4
    max_in_plane_deviation: float = math.radians(10.0)
    max_tilt: float = math.radians(10.0)
    seed: int = 1


@dataclass(frozen=True)
class UniformPosition:
    """Uniform centers over each fiber's feasible region."""


@dataclass(frozen=True)
class LayeredPosition:
    """Centers on ``layer_count`` evenly spaced layers with a uniform jitter
    of ``jitter_fraction`` layer spacings; each fiber records its layer as
    ``formation_layer``. ``axis`` None is the stack axis (generators.rs:263-300)."""

    layer_count: int
    axis: object = None
    jitter_fraction: float = 0.25


@dataclass(frozen=True)
class DensityGradientPosition:
    """Centers concentrated by a power law toward one side of ``axis``
    (generators.rs:302-330); ``axis`` None is the stack axis."""

    axis: object = None
    exponent: float = 1.0
    toward_high: bool = True


_DEFAULT_MATERIAL = Material("fiber", 0.025, 0.08)  # FiberPopulationSpec::default (population.rs:146-175)


def _range(value, name):
    if value is None:
        return None
    if isinstance(value, (int, float)):
        return (float(value), float(value))
    lo, hi = (float(v) for v in value)
    if lo > hi:
        raise ValueError(f"{name} range must be ordered as (min, max)")
    return (lo, hi)


@dataclass
class FiberPopulation:
    """Random fibers for ``generate_fiber_population`` (generators.rs:384-600).

    Unset, ``length`` is 16-24 mean diameters and ``curvature_amplitude``
    0-0.8 mean diameters; ``diameter`` defaults to the material's.
    """

    material: Material = field(default_factory=lambda: _DEFAULT_MATERIAL)
    count: int = 32
    segments_per_fiber: int = 8
    seed: int = 1
    length: object = None
    diameter: object = None
    curvature_amplitude: object = None
    nominal_parent_length: float | None = None
    orientation: object = field(default_factory=IsotropicOrientation)
    position: object = field(default_factory=UniformPosition)
    max_attempts_per_fiber: int = 128

    def copy(self) -> "FiberPopulation":
        return _dc_replace(self)

    def replace(self, **changes) -> "FiberPopulation":
        return _dc_replace(self, **changes)

    def _spec(self, cell: Cell, stack_axis: int | None = None) -> _native.PopulationSpec:
        """FiberPopulation.to_rust (generators.rs:570-600)."""
        stack = cell.stack_axis if stack_axis is None else stack_axis
        unit = [0.0, 0.0, 0.0]
        unit[stack] = 1.0
        spec = _native.PopulationSpec()
        diameter = _range(self.diameter, "diameter") or (self.material.diameter, self.material.diameter)
        mean_d = diameter[0] if diameter[0] == diameter[1] else 0.5 * (diameter[0] + diameter[1])
        length = _range(self.length, "length") or (16.0 * mean_d, 24.0 * mean_d)
        amplitude = _range(self.curvature_amplitude, "curvature_amplitude") or (0.0, 0.8 * mean_d)
        if self.material.is_oval:
            raise ValueError("oval sections are not supported by the hip backend")
        if isinstance(self.orientation, LayeredBiaxialOrientation) and not isinstance(self.position, LayeredPosition):
            raise ValueError("LayeredBiaxialOrientation picks a direction per layer, so it needs "
                             "position=LayeredPosition(...)")
        spec.count = int(self.count)
        spec.segments_per_fiber = int(self.segments_per_fiber)
        spec.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        spec.length_min, spec.length_max = length
        spec.radius_min, spec.radius_max = 0.5 * diameter[0], 0.5 * diameter[1]
        spec.amplitude_min, spec.amplitude_max = amplitude
        o = self.orientation
        if isinstance(o, IsotropicOrientation):
            spec.orientation = _native.ORIENT_ISOTROPIC
        elif isinstance(o, PlanarOrientation):
            spec.orientation = _native.ORIENT_PLANAR
            spec.orientation_vector[:] = _direction(o.normal)
            spec.orientation_angle = float(o.max_tilt)
        elif isinstance(o, AlignedOrientation):
            spec.orientation = _native.ORIENT_ALIGNED
            spec.orientation_vector[:] = _direction(o.axis)
            spec.orientation_angle = float(o.max_angle)
        elif isinstance(o, SchladitzOrientation):
            spec.orientation = _native.ORIENT_SCHLADITZ
            spec.orientation_vector[:] = _direction(o.axis)
            spec.orientation_beta = float(o.beta)
        elif isinstance(o, LayeredBiaxialOrientation):
            spec.orientation = _native.ORIENT_LAYERED_BIAXIAL
            spec.orientation_vector[:] = unit if o.normal is None else _direction(o.normal)
            spec.orientation_angle = float(o.max_tilt)
            spec.orientation_primary_fraction = float(o.primary_fraction)
            spec.orientation_cross_fraction = float(o.cross_fraction)
            spec.orientation_in_plane_deviation = float(o.max_in_plane_deviation)
            spec.orientation_layer_seed = int(o.seed) & 0xFFFFFFFFFFFFFFFF
        else:
            raise TypeError(f"unsupported orientation {o!r}")
        p = self.position
        if isinstance(p, UniformPosition):
            spec.position = _native.POSITION_UNIFORM
        elif isinstance(p, LayeredPosition):
            spec.position = _native.POSITION_LAYERED
            spec.position_axis = stack if p.axis is None else _AXES[p.axis]
            if int(p.layer_count) <= 0:
                raise ValueError("layer_count must be positive")
            spec.position_layers = int(p.layer_count)
            spec.position_jitter_fraction = float(p.jitter_fraction)
        elif isinstance(p, DensityGradientPosition):
            spec.position = _native.POSITION_GRADIENT
            spec.position_axis = stack if p.axis is None else _AXES[p.axis]
            spec.position_exponent = float(p.exponent)
            spec.position_toward_high = 1 if p.toward_high else 0
        else:
            raise TypeError("position must be UniformPosition, LayeredPosition, or DensityGradientPosition")
        spec.min_bend_radius = self.material.min_bend_radius or 0.0
        spec.max_attempts_per_fiber = int(self.max_attempts_per_fiber)
        spec.cell_origin[:] = cell.origin
        spec.cell_lengths[:] = cell.lengths
        spec.cell_periodic[:] = [1 if p else 0 for p in cell.periodic]
        return spec


@dataclass
class _Fiber:
    material: Material
    radius: float
    placed: np.ndarray      # (n, 3) float64
    intrinsic: np.ndarray   # (n, 3) float64
    formation_layer: int | None = None
    step: int = 0           # formation step, set by Recipe.insert


class FiberCollection:
    """Fibers with their rest (intrinsic) and placed centrelines."""

    def __init__(self, name: str = "fibers"):
        self.name = name
        self._fibers: list[_Fiber] = []

    def add_fiber(self, points, material: Material, rest_points=None, *, formation_layer: int | None = None) -> int:
        """collection.rs:276-307; `formation_layer` labels the fiber's ply."""
        placed = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        intrinsic = placed.copy() if rest_points is None else np.asarray(rest_points, dtype=np.float64).reshape(-1, 3)
        if placed.shape[0] < 2 or intrinsic.shape != placed.shape:
            raise ValueError("a fiber needs at least two points and matching rest points")
        if not (np.all(np.isfinite(placed)) and np.all(np.isfinite(intrinsic))):
            raise ValueError("centerline coordinates must be finite")
        layer = None if formation_layer is None else int(formation_layer)
        if layer is not None and not 0 <= layer < 0xFFFFFFFF:
            raise ValueError("formation_layer must be a u32 below 0xffffffff")
        self._fibers.append(_Fiber(material, material.radius, placed, intrinsic, layer))
        return len(self._fibers) - 1

    @classmethod
    def from_centerlines(cls, centerlines, material: Material, *, name: str = "collection",
                         formation_layer: int | None = None) -> "FiberCollection":
        """collection.rs:310-331."""
        out = cls(name)
        for line in centerlines:
            out.add_fiber(line, material, formation_layer=formation_layer)
        return out

    def layer_ids(self) -> list[int]:
        """The sorted formation-layer labels present (collection.rs:366-376)."""
        return sorted({f.formation_layer for f in self._fibers if f.formation_layer is not None})

    formation_layers = layer_ids

    def select_layer(self, layer: int, *, name: str | None = None, allow_empty: bool = False) -> "FiberCollection":
        """A copy of the fibers labelled `layer` (collection.rs:378-396)."""
        out = FiberCollection(name or f"{self.name} layer {layer}")
        out._fibers = [f for f in self._fibers if f.formation_layer == layer]
        if not out._fibers and not allow_empty:
            raise ValueError(f"collection contains no fibers in layer {layer}")
        return out

    def centerlines(self):
        return [f.placed.tolist() for f in self._fibers]

    def rest_centerlines(self):
        return [f.intrinsic.tolist() for f in self._fibers]

    def extend(self, other: "FiberCollection") -> None:
        self._fibers.extend(other._fibers)

    def __add__(self, other: "FiberCollection") -> "FiberCollection":
        out = FiberCollection(self.name)
        out._fibers = self._fibers + other._fibers
        return out

    def __len__(self) -> int:
        return len(self._fibers)


def generate_fiber_population(cell: Cell, population: FiberPopulation, *, name: str = "population") -> FiberCollection:
    """Places ``population``'s seeded random fibers in ``cell`` (population.rs:246-357)."""
    return _generate(cell, population, cell.stack_axis, name)


def _generate(cell: Cell, population: FiberPopulation, stack_axis: int, name: str) -> FiberCollection:
    spec = population._spec(cell, stack_axis)
    intrinsic, placed, radius, layers = _native.generate_population(spec)
    nv = population.segments_per_fiber + 1
    out = FiberCollection(name)
    for f in range(population.count):
        sl = slice(f * nv, (f + 1) * nv)
        layer = None if layers[f] == 0xFFFFFFFF else int(layers[f])
        out._fibers.append(_Fiber(population.material, float(radius[f]), placed[sl], intrinsic[sl], layer))
    return out


_BACKENDS = ("hip",)
_AGGREGATIONS = {"uniform_average": 0, "penetration_weighted": 1, "deepest_only": 2}


@dataclass
class RelaxationSettings:
    """Run-wide solver settings (settings.rs:250-609). Unset,
    ``penetration_tolerance`` is 1 % and ``max_step`` 10 % of the smallest
    fiber diameter in the recipe."""

    backend: str = "hip"
    motion_model: str = "flexible"
    pin_fiber_ends: bool = False
    penetration_tolerance: float | None = None
    force_full_iterations: bool = False
    correction_fraction: float = 0.8
    contact_aggregation: str = "uniform_average"
    stretch_stiffness: float = 0.35
    bend_stiffness: float = 0.03
    curvature_limit_stiffness: float = 1.0
    curvature_limit_safety_margin: float = 1.0e-3
    curvature_ratio_tolerance: float = 1.0e-5
    constraint_iterations: int = 2
    curvature_cleanup_sweeps: int = 4
    max_step: float | None = None
    max_iterations: int = 2000
    iterations_per_batch: int = 128
    cell_size_scale: float = 1.0
    neighbor_skin_scale: float = 2.0
    neighbor_capacity: int = 48

    def __post_init__(self):
        self._validate()

    def copy(self) -> "RelaxationSettings":
        return _dc_replace(self)

    def replace(self, **changes) -> "RelaxationSettings":
        return _dc_replace(self, **changes)

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k in self.__dataclass_fields__}

    def _validate(self) -> None:
        if self.backend not in _BACKENDS:
            raise ValueError(
                f"backend must be one of {list(_BACKENDS)}, got {self.backend!r} "
                "(this package builds only the HIP backend)"
            )
        if self.motion_model != "flexible":
            raise ValueError("the hip backend implements motion_model='flexible'")
        if self.contact_aggregation not in _AGGREGATIONS:
            raise ValueError(f"contact_aggregation must be one of {list(_AGGREGATIONS)}")
        for name in ("correction_fraction", "stretch_stiffness", "bend_stiffness", "curvature_limit_stiffness"):
            value = getattr(self, name)
            if not math.isfinite(value) or not (0.0 <= value <= 1.0) or (value == 0.0 and name == "correction_fraction"):
                raise ValueError(f"{name} must be finite and within its [0, 1] range")
        if not (math.isfinite(self.curvature_limit_safety_margin) and 0.0 <= self.curvature_limit_safety_margin < 0.1):
            raise ValueError("curvature_limit_safety_margin must be in [0, 0.1)")
        if not (math.isfinite(self.curvature_ratio_tolerance) and self.curvature_ratio_tolerance >= 0.0):
            raise ValueError("curvature_ratio_tolerance must be nonnegative and finite")
        if min(self.constraint_iterations, self.curvature_cleanup_sweeps, self.max_iterations,
               self.iterations_per_batch) <= 0:
            raise ValueError("iteration and sweep counts must be positive")
        if not (math.isfinite(self.cell_size_scale) and self.cell_size_scale >= 1.0):
            raise ValueError("cell_size_scale must be finite and at least 1")
        if not (math.isfinite(self.neighbor_skin_scale) and self.neighbor_skin_scale >= 0.0):
            raise ValueError("neighbor_skin_scale must be finite and non-negative")
        if self.neighbor_capacity < 1:
            raise ValueError("neighbor_capacity must be at least 1")
        if self.penetration_tolerance is not None and not (
            math.isfinite(self.penetration_tolerance) and self.penetration_tolerance >= 0.0
        ):
            raise ValueError("penetration_tolerance must be nonnegative and finite")
        if self.max_step is not None and not (math.isfinite(self.max_step) and self.max_step > 0.0):
            raise ValueError("max_step must be positive and finite")

    def to_config(self, smallest_diameter: float | None) -> _native.Config:
        """RelaxationConfig for a recipe whose thinnest fiber is ``smallest_diameter``
        (PyRelaxationSettings::to_rust_for, settings.rs:485-609)."""
        self._validate()
        c = _native.Config()
        tol = self.penetration_tolerance
        if tol is None:
            tol = 1.0e-4 if smallest_diameter is None else 0.01 * smallest_diameter
        step = self.max_step
        if step is None:
            step = 0.006 if smallest_diameter is None else 0.1 * smallest_diameter
        c.penetration_tolerance = tol
        c.correction_fraction = self.correction_fraction
        c.stretch_stiffness = self.stretch_stiffness
        c.bend_stiffness = self.bend_stiffness
        c.curvature_limit_stiffness = self.curvature_limit_stiffness
        c.curvature_limit_safety_margin = self.curvature_limit_safety_margin
        c.curvature_ratio_tolerance = self.curvature_ratio_tolerance
        c.max_step = step
        c.cell_size_scale = self.cell_size_scale
        c.neighbor_skin_scale = self.neighbor_skin_scale
        c.contact_aggregation = _AGGREGATIONS[self.contact_aggregation]
        c.constraint_iterations = self.constraint_iterations
        c.curvature_cleanup_sweeps = self.curvature_cleanup_sweeps
        c.force_full_iterations = 1 if self.force_full_iterations else 0
        c.neighbor_capacity = self.neighbor_capacity
        c.maximum_neighbor_list_bytes = 0
        return c


_EXHAUSTION = ("fail", "continue_if_hard_ok")


@dataclass
class SolvePolicy:
    """A relaxation gate of ``Recipe.solve`` (settings.rs:800-900,
    tangle_generate/src/recipe.rs:380-481): the kernels solve toward the
    targets, the stage is accepted when the measured residuals are within the
    max_* limits, and ``max_iterations`` (plus ``max_extra_iterations``) bound it."""

    name: str = "relaxation"
    target_penetration: float = 1.0e-4
    target_curvature_ratio: float = 1.00001
    max_penetration: float | None = None
    max_curvature_ratio: float | None = None
    hard_penetration: bool = True
    hard_curvature: bool = True
    max_iterations: int = 2000
    max_extra_iterations: int | None = None
    on_budget_exhausted: str = "fail"

    def __post_init__(self):
        if not (math.isfinite(self.target_penetration) and self.target_penetration >= 0.0):
            raise ValueError("target_penetration must be nonnegative and finite")
        if not (math.isfinite(self.target_curvature_ratio) and self.target_curvature_ratio >= 1.0):
            raise ValueError("target_curvature_ratio must be at least 1")
        if self.max_iterations <= 0:
            raise ValueError("max_iterations must be positive")
        if self.on_budget_exhausted not in _EXHAUSTION:
            raise ValueError(f"on_budget_exhausted must be one of {list(_EXHAUSTION)}")

    @property
    def penetration_limit(self) -> float:
        return self.target_penetration if self.max_penetration is None else self.max_penetration

    @property
    def curvature_limit(self) -> float:
        return self.target_curvature_ratio if self.max_curvature_ratio is None else self.max_curvature_ratio

    @property
    def extra_iterations(self) -> int:
        return self.max_iterations // 2 if self.max_extra_iterations is None else self.max_extra_iterations

    def accepts(self, penetration: float, curvature: float) -> bool:
        return penetration <= self.penetration_limit and curvature <= self.curvature_limit

    def hard_limits_hold(self, penetration: float, curvature: float) -> bool:
        return ((not self.hard_penetration or penetration <= self.penetration_limit)
                and (not self.hard_curvature or curvature <= self.curvature_limit))

    def copy(self) -> "SolvePolicy":
        return _dc_replace(self)

    def replace(self, **changes) -> "SolvePolicy":
        return _dc_replace(self, **changes)


# (correction_fraction, stretch, curvature_limit_stiffness, constraint_iterations,
# curvature_cleanup_sweeps); every preset uses deepest-only contact aggregation
# and no rest-chord bending (settings.rs:645-690).
_PRESETS = {
    "contact_first": (0.8, 0.05, 0.15, 1, 1),
    "curvature_cleanup": (0.05, 0.35, 1.0, 8, 16),
    "contact_cleanup": (0.8, 0.35, 0.75, 8, 16),
}


@dataclass
class RelaxationOverrides:
    """Per-stage replacements for RelaxationSettings fields, layered onto each
    batch of the stage (plugin.rs:902-942, 1069-1102)."""

    motion_model: str | None = None
    correction_fraction: float | None = None
    contact_aggregation: str | None = None
    stretch_stiffness: float | None = None
    bend_stiffness: float | None = None
    curvature_limit_stiffness: float | None = None
    constraint_iterations: int | None = None
    curvature_cleanup_sweeps: int | None = None

    @classmethod
    def preset(cls, name: str, **changes) -> "RelaxationOverrides":
        if name not in _PRESETS:
            raise ValueError(f"overrides preset must be one of {list(_PRESETS)}")
        cf, stretch, limit, iterations, sweeps = _PRESETS[name]
        base = cls(motion_model="flexible", correction_fraction=cf, contact_aggregation="deepest_only",
                   stretch_stiffness=stretch, bend_stiffness=0.0, curvature_limit_stiffness=limit,
                   constraint_iterations=iterations, curvature_cleanup_sweeps=sweeps)
        return _dc_replace(base, **changes)

    def copy(self) -> "RelaxationOverrides":
        return _dc_replace(self)

    def replace(self, **changes) -> "RelaxationOverrides":
        return _dc_replace(self, **changes)

    def apply(self, config: "_native.Config") -> "_native.Config":
        out = _native.Config.from_buffer_copy(config)
        if self.motion_model not in (None, "flexible"):
            raise ValueError("the hip backend implements motion_model='flexible'")
        if self.correction_fraction is not None:
            out.correction_fraction = self.correction_fraction
        if self.contact_aggregation is not None:
            out.contact_aggregation = _AGGREGATIONS[self.contact_aggregation]
        if self.stretch_stiffness is not None:
            out.stretch_stiffness = self.stretch_stiffness
        if self.bend_stiffness is not None:
            out.bend_stiffness = self.bend_stiffness
        if self.curvature_limit_stiffness is not None:
            out.curvature_limit_stiffness = self.curvature_limit_stiffness
        if self.constraint_iterations is not None:
            out.constraint_iterations = self.constraint_iterations
        if self.curvature_cleanup_sweeps is not None:
            out.curvature_cleanup_sweeps = self.curvature_cleanup_sweeps
        return out


class RecipeError(RuntimeError):
    """A recipe operation failed (tangle_python/src/recipe.rs:705-722)."""

    def __init__(self, operation_index: int, operation: str, iteration: int, reason: str):
        super().__init__(f"operation {operation_index} ({operation}) failed at iteration {iteration}: {reason}")
        self.operation_index = operation_index
        self.operation = operation
        self.iteration = iteration
        self.reason = reason


@dataclass
class RunResult:
    """The relaxed fibers and run statistics of ``Recipe.run``."""

    fiber_count: int
    iterations: int
    converged: bool
    max_penetration: float
    max_curvature_ratio: float
    max_displacement: float
    active_segments: int
    active_vertices: int
    cell_count: int
    uploaded_bytes: int
    downloaded_bytes: int
    events: list
    warnings: list
    neighbor_list_rebuilds: int
    wall_seconds: float
    positions: np.ndarray
    vertex_counts: np.ndarray
    post_recipe_iterations: int = 0
    segment_iterations: int = 0  # active segments x iterations over the run
    needling: list = field(default_factory=list)  # (layer, selected vertices, depth) per needle command
    compactions: list = field(default_factory=list)  # one report per compaction operation
    junctions: list = field(default_factory=list)  # tangle_amd.junctions.Junction records
    junction_captures: list = field(default_factory=list)  # one report per junction capture
    compaction_steps: list = field(default_factory=list)  # one report per accepted compaction step
    resumed: bool = False
    resumed_iteration: int | None = None
    checkpoint_saves: int = 0
    last_checkpoint_iteration: int | None = None
    last_checkpoint_bytes: int | None = None
    _fibers: list = field(default=None, repr=False)  # the recipe's fibers, in packed order
    _cell: object = field(default=None, repr=False)
    _bend: list = field(default=None, repr=False)  # per-fiber bend radius at the end of the run

    def write_ovito(self, path, *, view_script_path=None, session_path=None, coloring: str = "fiber",
                    bond_radius_ratio: float | None = 1.0) -> None:
        """Writes the relaxed fibers as one OVITO frame, one spherocylinder per
        segment, and optionally the script that styles it and saves a session
        (RunResult.write_ovito, crates/tangle_python/src/recipe.rs:1005-1030;
        tangle_export/src/ovito.rs). Junctions captured by the run are drawn
        as binder bonds of `bond_radius_ratio` times the thinner fiber's
        radius (None: no bonds; tangle_export/src/bridges.rs)."""
        from . import export
        if coloring not in export.COLORINGS:
            raise ValueError(f"coloring must be one of {', '.join(export.COLORINGS)}")
        if self._fibers is None:
            raise RuntimeError("this result does not carry its fibers")
        xyz = self.positions.reshape(-1, 3).astype(np.float64)
        rows, start = [], 0
        bend = self._bend or [f.material.min_bend_radius for f in self._fibers]
        for index, (fiber, n) in enumerate(zip(self._fibers, self.vertex_counts)):
            rows.append((index + 1, fiber.radius, xyz[start:start + n], fiber.intrinsic, bend[index]))
            start += n
        cell = self._cell
        lower = [float(o) for o in cell.origin]
        upper = [float(o) + float(l) for o, l in zip(cell.origin, cell.lengths)]
        export.write_ovito_frame(path, rows, lower, upper, cell.periodic, self.iterations,
                                 junctions=self.junctions, bond_radius_ratio=bond_radius_ratio)
        if view_script_path is not None:
            export.write_ovito_view_script(view_script_path, path, session_path=session_path, coloring=coloring,
                                           bond_radius_ratio=bond_radius_ratio)

    @property
    def junction_count(self) -> int:
        """Persistent junctions captured by the run (junction_count,
        tangle_python/src/recipe.rs:1001-1004)."""
        return len(self.junctions)

    def centerlines(self):
        out, start = [], 0
        xyz = self.positions.reshape(-1, 3).astype(np.float64)
        for n in self.vertex_counts:
            out.append(xyz[start:start + n].tolist())
            start += n
        return out

    def __repr__(self) -> str:
        return (f"RunResult(fibers={self.fiber_count}, iterations={self.iterations}, converged={self.converged}, "
                f"max_penetration={self.max_penetration:.3e}, max_curvature_ratio={self.max_curvature_ratio:.6f})")


def _status_json(st) -> dict:
    return {name: (float(v) if isinstance(v, float) else int(v)) for name, v in st.as_dict().items()}


def _status_from_json(d: dict) -> "_native.BatchStatus":
    st = _native.BatchStatus()
    for name, _ in st._fields_:
        setattr(st, name, d[name])
    return st


class _RunState:
    """Batch controller of one Recipe.run: sizes each device batch, layers
    stage settings onto it and applies the stage's gate (plugin.rs:1041-1181,
    tangle_generate/src/recipe.rs:1311-1518)."""

    def __init__(self, settings: RelaxationSettings, config: "_native.Config", warm_start=None):
        self.settings = settings
        self.config = config
        self.total = 0
        self.status = None
        self.status_current = False  # the status describes the current positions
        self.compactions = []        # CompactionReport per compaction operation
        self.junctions = []          # persistent junctions, in creation order
        self.junction_captures = []  # JunctionCaptureReport per capture
        self.compaction_steps = []   # CompactionStepReport per accepted step
        self.converged = False
        self.events: list[str] = []
        self.warnings: list[str] = []
        self.uploaded = 0
        self.downloaded = 0
        self.warm_start = warm_start
        self.last_proposal = 0
        self.history_marked = False
        self.proposals_kept = 0
        self.proposals_reverted = 0
        self.post_recipe_iterations = 0
        self.needling: list = []
        self.active_segments = 0
        self.segment_iterations = 0  # active segments x iterations, summed over batches
        # operation in progress and the iteration it started at (checkpoints)
        self.op_index = 0
        self.op_begin = 0
        self.checkpoint_interval = 0
        self.checkpoint_hook = None
        self.last_saved = None
        # a compaction in progress ({"index", "a"}: its controller state) for
        # checkpoints, and one restored from a checkpoint for its operation
        self.compaction_inflight = None
        self.resume_compaction = None

    def _propose(self, world, config, room):
        """Learned warm start: every `interval` iterations of the first
        `active_iterations`, move the vertices by the network's prediction,
        run `settle` plain iterations, and revert the move (keeping the
        iteration count) if the normalised residual max(penetration /
        tolerance, curvature excess / tolerance) rose above `tolerance` times
        its value before the move. Plain PBD still runs every later pass.
        A due proposal whose settle iterations would not leave one iteration
        of the `room` the stage and the run have left is skipped; returns
        False then, True otherwise."""
        ws = self.warm_start
        if ws is None or self.total >= ws.active_iterations:
            return True
        from . import learned
        if not self.history_marked:
            # the first interval only provides the motion history
            learned.mark_history(world)
            self.history_marked = True
            self.last_proposal = self.total
            return True
        if self.total - self.last_proposal < ws.interval:
            return True
        if ws.settle >= room:
            return False
        self.last_proposal = self.total
        pen_tol = max(config.penetration_tolerance, 1e-30)
        curv_tol = max(config.curvature_ratio_tolerance, 1e-30)

        def residual(st):
            return max(st.max_penetration / pen_tol, (st.max_curvature_ratio - 1.0) / curv_tol)

        # a batch's last pass only measures, so the status of the batch that
        # brought the positions here describes them; it is reused rather
        # than measured again
        prior = self.status if self.status is not None and self.status_current else learned.measure(world, config)
        before = residual(prior)
        learned.propose(world, ws.model, ws.scale)
        if ws.settle > 0:
            settle = _native.Config.from_buffer_copy(config)
            settle.force_full_iterations = 1
            # no checkpoint inside the settle: its positions may be reverted;
            # a save that falls due is taken after the keep/revert decision
            after = residual(self._batch(world, settle, ws.settle, checkpoint=False))
        else:
            after = residual(learned.measure(world, config))
        if after > ws.tolerance * before:
            learned.revert(world)
            self.proposals_reverted += 1
            # the reverted positions are the ones `prior` describes
            self.status = prior
            self.converged = False
        else:
            self.proposals_kept += 1
        learned.mark_history(world)
        self.last_proposal = self.total
        self._checkpoint_if_due()
        return True

    def _batch(self, world, config, iterations, checkpoint=True):
        status = world.run_batch(config, iterations)
        self.segment_iterations += self.active_segments * int(status.batch_iterations)
        self.downloaded += 11 * 4 + 3 * 4
        self.total = int(status.total_iterations)
        self.status = status
        self.status_current = True
        self.converged = bool(status.converged)
        if checkpoint:
            self._checkpoint_if_due()
        return status

    def _checkpoint_if_due(self):
        """Saves when the iteration count has crossed an interval boundary
        since the last save."""
        if self.checkpoint_hook is not None and self.total > 0 and self.last_saved != self.total:
            n = self.checkpoint_interval
            crossed = self.total >= n if self.last_saved is None else self.total // n > self.last_saved // n
            if crossed:
                self.save_checkpoint()

    def save_checkpoint(self, failure: str | None = None):
        """Saves the world and the recipe's progress (save_checkpoint,
        tangle_checkpoint/src/lib.rs:318-375): when an interval boundary is
        crossed, when the run ends and when it fails."""
        if self.checkpoint_hook is None or (self.last_saved == self.total and failure is None):
            return
        self.checkpoint_hook(failure)
        self.last_saved = self.total

    def compaction_state(self):
        """The in-flight compaction's controller state as JSON-ready data and
        its accepted device state (None when no compaction is running):
        (state, accepted checkpoint or None)."""
        inflight = self.compaction_inflight
        if inflight is None or inflight["index"] != self.op_index:
            return None, None
        a = inflight["a"]
        m = a["metrics"]
        st = a["accepted_status"]
        state = {
            "index": inflight["index"], "next": float(a["next"]), "started": a["started"],
            "windows": int(a["windows"]), "steps": int(a["steps"]), "last": float(a["last"]),
            "previous": [float(v) for v in a["previous"]], "work": [float(w) for w in a["work"]],
            "baseline": bool(a["baseline"]),
            "metrics": None if m is None else {
                "pressure": [float(v) for v in m.pressure], "contact_energy": float(m.contact_energy),
                "stretch_energy": float(m.stretch_energy), "bending_energy": float(m.bending_energy),
                "total_energy": float(m.total_energy),
                "face_pressures": [[float(v) for v in pair] for pair in m.face_pressures]},
            "accepted_status": None if st is None else _status_json(st),
        }
        return state, a["accepted"]

    @staticmethod
    def compaction_from_state(state: dict, accepted) -> dict:
        """The controller state `compaction_state` saved, rebuilt."""
        f32 = np.float32
        m = state["metrics"]
        metrics = None if m is None else _native.CompactionMetrics(
            tuple(m["pressure"]), m["contact_energy"], m["stretch_energy"], m["bending_energy"], m["total_energy"],
            tuple(tuple(pair) for pair in m["face_pressures"]))
        a = dict(next=f32(state["next"]), started=state["started"], windows=int(state["windows"]),
                 steps=int(state["steps"]), last=f32(state["last"]), previous=list(state["previous"]),
                 work=[f32(w) for w in state["work"]], metrics=metrics, baseline=bool(state["baseline"]),
                 accepted=accepted,
                 accepted_status=None if state["accepted_status"] is None else _status_from_json(
                     state["accepted_status"]))
        return {"index": int(state["index"]), "a": a}

    def _proposal_window(self, world, config, batch, room):
        """Proposes when one is due and fits in `room` iterations, and
        shortens the batch so that the next proposal falls on a batch
        boundary; the caller recounts its budget afterwards, since the
        proposal's settle iterations count."""
        ws = self.warm_start
        if ws is None or self.total >= ws.active_iterations:
            return batch
        if not self._propose(world, config, room):
            return batch
        return max(1, min(batch, self.last_proposal + ws.interval - self.total))

    def _budget(self, index, description):
        if self.total >= self.settings.max_iterations:
            raise RecipeError(index, description, self.total,
                              f"iteration budget {self.settings.max_iterations} exhausted")
        return self.settings.max_iterations - self.total

    def _residuals(self):
        s = self.status
        return (f"penetration {s.max_penetration:.3e}, curvature ratio {s.max_curvature_ratio:.6f}"
                if s is not None else "no measurement")

    def relax(self, world, index, description, kind, amount, begin=None):
        begin = self.total if begin is None else begin
        config = _native.Config.from_buffer_copy(self.config)
        force_full = kind == "relax_for"
        config.force_full_iterations = 1 if (force_full or self.settings.force_full_iterations) else 0
        self.converged = False
        while True:
            completed = self.total - begin
            if kind == "relax_until_converged" and self.converged:
                self.events.append(f"relax to convergence in {completed} iterations ({self._residuals()})")
                return
            if completed >= amount:
                if force_full:
                    self.events.append(f"relax for {amount} iterations")
                    return
                raise RecipeError(index, description, self.total,
                                  f"failed to converge in {amount} iterations: {self._residuals()}")
            remaining = self._budget(index, description)
            room = min(remaining, amount - completed)
            batch = min(room, self.settings.iterations_per_batch)
            batch = self._proposal_window(world, config, batch, room)
            room = min(self.settings.max_iterations - self.total, amount - (self.total - begin))
            if room <= 0:
                continue
            status = self._batch(world, config, min(batch, room))
            if status.cell_list_overflow:
                raise RecipeError(index, description, self.total, "cell list overflow")

    def settle(self, world, index, description, tolerance: float, maximum: int, begin=None):
        """RelaxUntilTargetsReached (recipe.rs:1521-1596): full batches until
        the largest layer-centre or needle residual is within `tolerance`."""
        begin = self.total if begin is None else begin
        config = _native.Config.from_buffer_copy(self.config)
        config.force_full_iterations = 1
        tol = np.float32(tolerance)
        while True:
            completed = self.total - begin
            residuals = [e for e in world.formation_target_error() if e is not None]
            if not residuals:
                self.events.append("no active formation target to advance")
                return
            worst = max(residuals)
            if np.float32(worst) <= tol:
                self.events.append(f"advance formation targets in {completed} iterations (residual {worst:.3e})")
                return
            if completed >= maximum:
                raise RecipeError(index, description, self.total,
                                  f"formation targets remained at residual {worst:.3e} after {maximum} iterations "
                                  f"(tolerance {tolerance:.3e})")
            remaining = self._budget(index, description)
            batch = min(remaining, self.settings.iterations_per_batch, maximum - completed)
            status = self._batch(world, config, batch)
            if status.cell_list_overflow:
                raise RecipeError(index, description, self.total, "cell list overflow")

    def finish(self, world, begin=None):
        """After the last operation the targets are released and relaxation
        continues, unconstrained, until the run's tolerances hold or its
        iteration budget is spent (recipe.rs:893-915, plugin.rs:1041-1181);
        not converging here is not an error."""
        world.clear_layer_targets()
        world.clear_vertex_targets()
        begin = self.total if begin is None else begin
        config = _native.Config.from_buffer_copy(self.config)
        config.force_full_iterations = 1 if self.settings.force_full_iterations else 0
        self.converged = False
        while self.total < self.settings.max_iterations:
            room = self.settings.max_iterations - self.total
            batch = min(room, self.settings.iterations_per_batch)
            batch = self._proposal_window(world, config, batch, room)
            room = self.settings.max_iterations - self.total
            if room <= 0:
                break
            status = self._batch(world, config, min(batch, room))
            if status.cell_list_overflow or self.converged:
                break
        self.post_recipe_iterations = self.total - begin

    def solve(self, world, index, description, policy: "SolvePolicy", overrides, begin=None):
        begin = self.total if begin is None else begin
        config = _native.Config.from_buffer_copy(self.config)
        if overrides is not None:
            config = overrides.apply(config)
        config.force_full_iterations = 0
        config.penetration_tolerance = policy.target_penetration
        config.curvature_ratio_tolerance = max(policy.target_curvature_ratio - 1.0, 0.0)
        measured = False
        while True:
            completed = self.total - begin
            s = self.status
            if measured and s is not None and policy.accepts(s.max_penetration, s.max_curvature_ratio):
                self.events.append(f"{policy.name} accepted in {completed} iterations ({self._residuals()})")
                return
            if completed < policy.max_iterations:
                remaining = policy.max_iterations - completed
            elif (policy.on_budget_exhausted == "continue_if_hard_ok" and s is not None
                  and policy.hard_limits_hold(s.max_penetration, s.max_curvature_ratio)):
                self.warnings.append(f"{policy.name} exhausted {policy.max_iterations} iterations with deferred "
                                     f"soft limits: {self._residuals()}")
                return
            elif completed < policy.max_iterations + policy.extra_iterations and \
                    self.settings.max_iterations - self.total > 1:
                remaining = min(policy.max_iterations + policy.extra_iterations - completed,
                                self.settings.max_iterations - self.total - 1)
            else:
                raise RecipeError(index, description, self.total,
                                  f"{policy.name} failed after {completed} iterations: {self._residuals()}")
            run_remaining = self._budget(index, description)
            room = min(run_remaining, remaining)
            batch = max(1, min(room, self.settings.iterations_per_batch))
            before = self.total
            batch = self._proposal_window(world, config, batch, room)
            room -= self.total - before
            if room <= 0:
                continue
            status = self._batch(world, config, min(batch, room))
            measured = True
            if status.cell_list_overflow:
                raise RecipeError(index, description, self.total, "cell list overflow")

    def capture_junctions(self, world, index, description, policy, ctx):
        """One junction capture: device candidates, then the host policy
        (capture_junctions, tangle_generate/src/recipe.rs:1660-1671)."""
        from . import junctions as jn
        capture = world.capture_contact_arrays(np.float32(policy.max_surface_gap), int(policy.candidate_capacity))
        self.downloaded += 24 * len(capture[2]) + 8
        try:
            report = jn.form_junctions(self.junctions, ctx["fibers"], ctx["counts"], ctx["intrinsic"], capture,
                                       policy)
        except OverflowError as err:
            raise RecipeError(index, description, self.total, str(err)) from None
        self.junction_captures.append(report)
        return report

    def _hold(self, world, config, index, description, remaining):
        """One relaxation batch of at most `remaining` iterations, converging
        early when the run's tolerances hold (hold_for_relaxation with
        force_full_batch off, compaction_control.rs:428-432; the batch size
        rule of plugin.rs:1050-1075)."""
        run_remaining = self._budget(index, description)
        batch = max(1, min(run_remaining, self.settings.iterations_per_batch, max(1, remaining)))
        status = self._batch(world, config, batch)
        if status.cell_list_overflow:
            raise RecipeError(index, description, self.total, "cell list overflow")
        return status

    def compact(self, ctx, index, description, spec, overrides):
        """Closed-loop compaction (advance_compaction and its helpers,
        tangle_generate/src/recipe/compaction_control.rs:5-432): an admissible
        baseline, then compact-relax-measure steps with adaptive strain,
        rollback to the last accepted state when a window does not relax, and
        a stop at the target, a guard or a jam. A checkpoint taken inside the
        operation holds the controller's state (`compaction_state`), and a
        resume continues the operation from it, as the reference's
        checkpoints carry the recipe's compaction state."""
        f32 = np.float32
        world = ctx["world"]
        config = _native.Config.from_buffer_copy(self.config)
        if overrides is not None:
            config = overrides.apply(config)
        config.force_full_iterations = 0
        model = _native.CompactionEnergyModel(spec.contact_energy_stiffness, spec.stretch_energy_stiffness,
                                              spec.bending_energy_stiffness)
        cf = float(config.correction_fraction)
        state = ctx["state"]
        resumed = self.resume_compaction if (self.resume_compaction or {}).get("index") == index else None
        self.resume_compaction = None
        if resumed is not None:
            # the controller state a checkpoint took inside this operation
            a = resumed["a"]
        else:
            metrics = world.compaction_metrics(cf, model)
            self.downloaded += 13 * 4
            a = dict(next=f32(spec.initial_log_strain), started=self.total, windows=0, steps=0, last=f32(0.0),
                     previous=list(state["cell"][1]), work=[f32(0.0)] * 3, metrics=metrics, baseline=True,
                     accepted=None, accepted_status=None)
        self.compaction_inflight = {"index": index, "a": a}
        try:
            self._compact_loop(ctx, index, description, spec, config, model, cf, a, resumed is not None)
        finally:
            self.compaction_inflight = None

    def _compact_loop(self, ctx, index, description, spec, config, model, cf, a, resumed):
        from . import compaction as cp
        f32 = np.float32
        world = ctx["world"]
        state = ctx["state"]
        active_step = state["active_step"]
        fibers = ctx["fibers"]
        formed = [f for f in fibers if active_step is None or f.step <= active_step]
        path = spec.weights_for(ctx["stack_axis"])
        R = int(spec.relax_iterations)
        solid = cp.nominal_fiber_volume(fibers, active_step)

        def lengths_now():
            return list(state["cell"][1])

        def finish(reason):
            lengths = lengths_now()
            fraction = solid / (lengths[0] * lengths[1] * lengths[2])
            m = a["metrics"]
            self.compactions.append({"operation": index, "reason": cp.reason_text(reason), "steps": a["steps"],
                                     "cell_lengths": lengths, "nominal_volume_fraction": fraction,
                                     "pressure": list(m.pressure), "total_energy": m.total_energy,
                                     "cumulative_work": [float(w) for w in a["work"]]})
            self.events.append(f"compact in {a['steps']} steps to nominal volume fraction {fraction:.4f} "
                               f"({cp.reason_text(reason)})")

        def accept(m):
            lengths = lengths_now()
            origin = state["cell"][0]
            accepted = a["accepted"]
            cur_lo = [f32(origin[k]) for k in range(3)]
            cur_hi = [f32(origin[k] + lengths[k]) for k in range(3)]
            motion, work = [[f32(0.0)] * 2 for _ in range(3)], [[f32(0.0)] * 2 for _ in range(3)]
            fp = m.face_pressures
            for k in range(3):
                area = lengths[(k + 1) % 3] * lengths[(k + 2) % 3]
                shortening = max(a["previous"][k] - lengths[k], 0.0)
                a["work"][k] = f32(a["work"][k] + f32(f32(m.pressure[k]) * f32(area * shortening)))
                lower_motion = f32(max(f32(cur_lo[k] - f32(accepted.cell_lower[k])), f32(0.0)))
                upper_motion = f32(max(f32(f32(accepted.cell_upper[k]) - cur_hi[k]), f32(0.0)))
                motion[k] = [lower_motion, upper_motion]
                work[k] = [f32(f32(f32(fp[k][0]) * f32(area)) * lower_motion),
                           f32(f32(f32(fp[k][1]) * f32(area)) * upper_motion)]
            a["steps"] += 1
            a["metrics"] = m
            st = self.status
            self.compaction_steps.append({
                "operation": index, "step": a["steps"], "iteration": self.total, "cell_lengths": lengths,
                "nominal_volume_fraction": solid / (lengths[0] * lengths[1] * lengths[2]),
                "log_volume_strain": float(a["last"]), "max_penetration": float(st.max_penetration),
                "max_bend_ratio": float(st.max_curvature_ratio), "pressure": list(m.pressure),
                "contact_energy": m.contact_energy, "stretch_energy": m.stretch_energy,
                "bending_energy": m.bending_energy, "total_energy": m.total_energy,
                "cumulative_work": [float(w) for w in a["work"]]})
            readings = []
            for k, name in enumerate("xyz"):
                if any(v > 0.0 for v in fp[k]) or any(v > 0.0 for v in work[k]):
                    readings.append(
                        f"{name}- p={cp.rust_exp3(fp[k][0])} dx={cp.rust_exp3(motion[k][0])} "
                        f"dW={cp.rust_exp3(work[k][0])}, {name}+ p={cp.rust_exp3(fp[k][1])} "
                        f"dx={cp.rust_exp3(motion[k][1])} dW={cp.rust_exp3(work[k][1])}")
            if readings:
                self.events.append(f"compaction step {a['steps']} wall pressure/motion/work: "
                                   f"{'; '.join(readings)}")

        def rollback():
            ck = a["accepted"]
            ck.total_iterations = self.total
            new_world = ctx["restore"](ck)
            state["cell"] = ([float(v) for v in ck.cell_lower],
                             [float(f32(f32(ck.cell_upper[k]) - f32(ck.cell_lower[k]))) for k in range(3)])
            self.status = a["accepted_status"]
            self.status_current = True
            return new_world

        # an unconstrained admissible baseline first (a resumed operation
        # continues where its checkpoint left it)
        if not resumed:
            self.converged = False
            self._hold(world, config, index, description, R)
        while True:
            if a["started"] is not None:
                elapsed = self.total - a["started"]
                if not self.converged and elapsed < R:
                    self._hold(world, config, index, description, R - elapsed)
                    continue
                if not self.converged:
                    a["windows"] += 1
                    if a["baseline"] and a["windows"] < spec.max_relax_windows:
                        a["started"] = self.total
                        self.converged = False
                        self._hold(world, config, index, description, R)
                        continue
                    if a["baseline"]:
                        a["metrics"] = world.compaction_metrics(cf, model)
                        self.downloaded += 13 * 4
                        finish(("GuardReached", "initial formation state did not relax (penalty energy "
                                f"{cp.rust_exp3(a['metrics'].total_energy)})"))
                        return
                    world = rollback()
                    attempted = a["next"]
                    reduced = f32(max(f32(attempted * f32(spec.shrink_factor)), f32(spec.min_log_strain)))
                    at_minimum = attempted <= f32(f32(spec.min_log_strain) * f32(1.0 + 16.0 * cp.F32_EPS))
                    if at_minimum or a["windows"] >= spec.max_relax_windows:
                        finish("Jammed")
                        return
                    a["next"] = reduced
                    a["started"] = None
                    continue
                m = world.compaction_metrics(cf, model)
                self.downloaded += 13 * 4
                a["metrics"] = m
                if a["baseline"]:
                    a["baseline"] = False
                    a["started"] = None
                    a["windows"] = 0
                    a["accepted"] = world.checkpoint()
                    a["accepted_status"] = self.status
                    continue
                accept(m)
                a["accepted"] = world.checkpoint()
                a["accepted_status"] = self.status
                a["started"] = None
                a["windows"] = 0
                if elapsed <= R // 2:
                    a["next"] = f32(min(f32(a["next"] * f32(spec.growth_factor)), f32(spec.max_log_strain)))
                else:
                    a["next"] = f32(max(f32(a["next"] * f32(spec.shrink_factor)), f32(spec.min_log_strain)))
            lengths = lengths_now()
            fraction = solid / (lengths[0] * lengths[1] * lengths[2])
            m = a["metrics"]
            st = self.status
            guard = cp.guard_reached(spec, st.max_penetration, st.max_curvature_ratio, m.pressure, m.total_energy,
                                     a["steps"])
            if guard is not None:
                finish(("GuardReached", guard))
                return
            if cp.target_reached(spec.target, spec.target_tolerance, lengths, fraction, m.pressure, m.total_energy):
                finish("TargetReached")
                return
            weights, remaining = cp.strain_direction_and_remaining(spec.target, path, lengths, fraction, m.pressure)
            requested = a["next"] if remaining is None else f32(min(a["next"], remaining))
            requested = f32(max(requested, f32(cp.F32_EPS)))
            if not formed:
                raise RecipeError(index, description, self.total, "compaction needs active fibers")
            min_diameter = min(2.0 * f.radius for f in formed)
            requested = cp.diameter_limited_log_strain(requested, weights, lengths, min_diameter,
                                                       spec.max_shortening_over_min_diameter)
            minimum_length = 2.001 * max(f.radius for f in formed)
            new_lengths = [max(lengths[k] * math.exp(-float(f32(requested * weights[k]))), minimum_length)
                           for k in range(3)]
            applied = f32(sum(math.log(lengths[k] / new_lengths[k]) for k in range(3)))
            if applied <= f32(cp.F32_EPS):
                finish(("GuardReached", "cell reached the minimum fiber-diameter extent"))
                return
            old_lower, _ = world.cell_bounds()
            anchor = [f32(v) for v in spec.cell_anchor]
            if spec.balance_opposing_faces:
                a["metrics"] = world.compaction_metrics(cf, model)
                self.downloaded += 13 * 4
                fp = a["metrics"].face_pressures
                floor = f32(spec.face_pressure_floor)
                strength = f32(spec.face_balance_strength)
                for k in range(3):
                    if weights[k] > 0.0 and not ctx["periodic"][k]:
                        lower_ease = f32(f32(1.0) / f32(f32(fp[k][0]) + floor))
                        upper_ease = f32(f32(1.0) / f32(f32(fp[k][1]) + floor))
                        adaptive = f32(lower_ease / f32(lower_ease + upper_ease))
                        base = f32(spec.cell_anchor[k])
                        anchor[k] = f32(base + f32(strength * f32(adaptive - base)))
            new_lower = [f32(f32(old_lower[k]) + f32(anchor[k] * f32(f32(lengths[k]) - f32(new_lengths[k]))))
                         for k in range(3)]
            new_upper = [f32(new_lower[k] + f32(new_lengths[k])) for k in range(3)]
            try:
                world.compact_cell(new_lower, new_upper, spec.kinematics)
            except (ValueError, _native.NativeError) as err:
                raise RecipeError(index, description, self.total, f"compaction step rejected: {err}") from None
            state["cell"] = ([float(v) for v in new_lower],
                             [float(f32(new_upper[k] - new_lower[k])) for k in range(3)])
            self.status_current = False
            a["previous"] = lengths
            a["last"] = applied
            a["started"] = self.total
            self.converged = False
            self._hold(world, config, index, description, R)


class HeldTargets:
    """Returned by operations that hold fibers on targets
    (crates/tangle_python/src/recipe.rs:148-195). As a context manager it
    records the matching release when the block ends; releasing layer
    placement releases every held layer."""

    def __init__(self, recipe: "Recipe", kind: str):
        self._recipe = recipe
        self._kind = kind

    def __enter__(self) -> "Recipe":
        return self._recipe

    def __exit__(self, *exc) -> bool:
        if self._kind == "layers":
            self._recipe.release_layer_placement()
        else:
            self._recipe.release_needles()
        return False

    def __repr__(self) -> str:
        return f"HeldTargets({'layer placement' if self._kind == 'layers' else 'needles'})"


def _positive(value, name):
    value = float(value)
    if not (math.isfinite(value) and value > 0.0):
        raise ValueError(f"{name} must be positive and finite")
    return value


def _unit_fraction(value, name):
    value = float(value)
    if not (math.isfinite(value) and 0.0 < value <= 1.0):
        raise ValueError(f"{name} must lie in (0, 1]")
    return value


def _curvature_limits(bend_radii) -> np.ndarray:
    """Per-fiber maximum curvature, f32 of 1 / radius, 0 without a limit
    (FiberBendLimit::maximum_curvature as the recipe uploads it)."""
    return np.array([0.0 if r is None else np.float32(1.0 / r) for r in bend_radii], dtype=np.float32)


def _active_capsule_bounds(fibers, positions, starts, active, padding):
    """Bounds of the active segments' capsules, each endpoint padded by its
    radius plus `padding`, in f32 (active_capsule_bounds,
    tangle_generate/src/recipe/compaction_control.rs:446-474); None when no
    segment is active."""
    lower = np.full(3, np.inf, dtype=np.float32)
    upper = np.full(3, -np.inf, dtype=np.float32)
    found = False
    pad = np.float32(padding)
    for index, fiber in enumerate(fibers):
        n = fiber.placed.shape[0]
        if not active[index] or n < 2:
            continue
        found = True
        radius = np.float32(np.float32(fiber.radius) + pad)
        pts = positions[int(starts[index]):int(starts[index]) + n]
        lower = np.minimum(lower, (pts - radius).min(axis=0))
        upper = np.maximum(upper, (pts + radius).max(axis=0))
    if not found:
        return None, None
    return [np.float32(v) for v in lower], [np.float32(v) for v in upper]


class Recipe:
    """An ordered list of formation operations on one cell; ``insert`` adds
    fibers at once, everything else runs in order inside ``run``
    (crates/tangle_python/src/recipe.rs:197-790). Every ``insert`` is a new
    formation step: its fibers join the relaxation when the recipe reaches
    it."""

    def __init__(self, cell: Cell, *, stack_axis=None):
        if not isinstance(cell, Cell):
            raise TypeError("Recipe needs a Cell")
        self._cell = cell
        self._fibers: list[_Fiber] = []
        self._operations: list[tuple] = []
        self._descriptions: list[str] = []
        self._next_step = 0
        self._stack_axis = cell.stack_axis if stack_axis is None else _AXES[stack_axis]

    @property
    def stack_axis(self) -> int:
        return self._stack_axis

    def _push(self, operation: tuple, description: str) -> None:
        self._operations.append(operation)
        self._descriptions.append(description)

    def insert(self, collection, *, name: str | None = None):
        """Adds the fibers as the next formation step (recipe.rs:250-298); a
        FiberPopulation is generated in this recipe's cell first, stacked
        along the recipe's axis."""
        if isinstance(collection, FiberPopulation):
            collection = _generate(self._cell, collection, self._stack_axis, "fiber population")
        if not isinstance(collection, FiberCollection):
            raise TypeError("insert takes a FiberCollection or a FiberPopulation")
        if len(collection) == 0:
            raise ValueError("cannot insert an empty fiber collection")
        step = self._next_step
        self._next_step += 1
        self._fibers.extend(_dc_replace(f, step=step) for f in collection._fibers)
        label = name or collection.name
        self._push(("activate", step), f"insert {label!r} ({len(collection)} fibers) as formation step {step}")
        return len(collection)

    def relax_until_converged(self, *, max_iterations: int = 2000) -> None:
        if max_iterations <= 0:
            raise ValueError("max_iterations must be positive")
        self._push(("relax_until_converged", int(max_iterations)),
                   f"relax until converged (up to {max_iterations} iterations)")

    def relax_for(self, iterations: int) -> None:
        if iterations <= 0:
            raise ValueError("iterations must be positive")
        self._push(("relax_for", int(iterations)), f"relax for exactly {iterations} iterations")

    def settle_targets(self, *, tolerance: float, max_iterations: int) -> None:
        """Relaxes until held layers and needles are within `tolerance` of
        their targets (recipe.rs:327-343)."""
        tolerance = float(tolerance)
        if not (math.isfinite(tolerance) and tolerance >= 0.0):
            raise ValueError("tolerance must be non-negative and finite")
        if max_iterations <= 0:
            raise ValueError("max_iterations must be positive")
        self._push(("settle_targets", (tolerance, int(max_iterations))),
                   f"settle held targets to tolerance {tolerance}")

    def solve(self, policy: SolvePolicy, overrides: RelaxationOverrides | None = None) -> None:
        """Relax under `policy`'s targets and gate, with `overrides` layered
        onto every batch of the stage (recipe.rs:1398-1518)."""
        if not isinstance(policy, SolvePolicy):
            raise TypeError("solve takes a SolvePolicy")
        if overrides is not None and not isinstance(overrides, RelaxationOverrides):
            raise TypeError("overrides must be RelaxationOverrides")
        self._push(("solve", (policy, overrides)), f"solve {policy.name!r}")

    def scale_layer_spacing(self, factor: float, *, stiffness: float = 0.25,
                            max_translation: float = 1.0e-5) -> HeldTargets:
        """Moves every layer toward `factor` times its initial spacing from the
        cell centre and holds it there (recipe.rs:395-416, 1027-1053)."""
        factor = _positive(factor, "factor")
        stiffness = _unit_fraction(stiffness, "stiffness")
        max_translation = _positive(max_translation, "max_translation")
        self._push(("scale_layers", (factor, stiffness, max_translation)), f"scale layer spacing by {factor}")
        return HeldTargets(self, "layers")

    def place_layer_above(self, layer: int, *, gap: float, stiffness: float = 0.25,
                          max_translation: float = 1.0e-5) -> HeldTargets:
        """Holds the fibers of formation layer `layer` at `gap` above layer
        `layer - 1`, leaving the layers below free (recipe.rs:418-444,
        1055-1089)."""
        layer = int(layer)
        if layer <= 0:
            raise ValueError("layer must be greater than zero")
        gap = _positive(gap, "gap")
        stiffness = _unit_fraction(stiffness, "stiffness")
        max_translation = _positive(max_translation, "max_translation")
        self._push(("place_layer", (layer, gap, stiffness, max_translation)),
                   f"place layer {layer} above layer {layer - 1}")
        return HeldTargets(self, "layers")

    def release_layer_placement(self) -> None:
        self._push(("release_layers", None), "release layer placement")

    def needle_layer(self, layer: int, *, footprint, depth: float, min_fiber_diameter: float | None = None,
                     stiffness: float = 0.25, max_translation: float = 1.0e-5,
                     max_translation_over_diameter: float = 0.25) -> HeldTargets:
        """Pulls one internal vertex of the selected fibers of formation layer
        `layer` by `depth` against the stack axis (recipe.rs:451-504,
        1091-1177)."""
        from . import formation
        layer = int(layer)
        depth = _positive(depth, "depth")
        if min_fiber_diameter is not None:
            min_fiber_diameter = _positive(min_fiber_diameter, "min_fiber_diameter")
        stiffness = _unit_fraction(stiffness, "stiffness")
        max_translation = _positive(max_translation, "max_translation")
        ratio = _positive(max_translation_over_diameter, "max_translation_over_diameter")
        if isinstance(footprint, formation.RandomFiberFraction):
            selection = footprint
            text = f"a random {footprint.fraction} of its fibers"
        elif isinstance(footprint, formation.CircularFootprint):
            if footprint.center is None:
                plane = [k for k in range(3) if k != self._stack_axis]
                origin = [np.float32(self._cell.origin[k]) for k in plane]
                extent = [np.float32(self._cell.lengths[k]) for k in plane]
                center = formation.random_footprint_center(footprint.seed or 0, layer, origin, extent)
            else:
                center = tuple(np.float32(c) for c in footprint.center)
            selection = (center, np.float32(footprint.diameter))
            text = f"a {footprint.diameter:.3e} circle at [{float(center[0]):.3e}, {float(center[1]):.3e}]"
        else:
            raise TypeError("footprint must be a CircularFootprint or RandomFiberFraction")
        needling = formation.Needling(layer, selection, min_fiber_diameter, depth, stiffness, max_translation, ratio)
        self._push(("needle", needling), f"needle layer {layer} through {text}")
        return HeldTargets(self, "needles")

    def release_needles(self) -> None:
        self._push(("release_needles", None), "release needles")

    def compact(self, settings, overrides: "RelaxationOverrides | None" = None) -> None:
        """Closed-loop compaction toward `settings.target`
        (tangle_python/src/recipe.rs:544-563; the controller is
        tangle_amd/compaction.py and _RunState.compact)."""
        from . import compaction as cp
        if not isinstance(settings, cp.CompactionSettings):
            raise TypeError("settings must be CompactionSettings")
        if overrides is not None and not isinstance(overrides, RelaxationOverrides):
            raise TypeError("overrides must be RelaxationOverrides")
        settings.validate()
        self._push(("compact", (settings.copy(), overrides)), f"compact toward {settings.target!r}")

    def capture_junctions(self, policy) -> None:
        """Promotes the current contacts to junctions under `policy`
        (tangle_python/src/recipe.rs:564-571; tangle_generate/src/recipe.rs:1597-1605)."""
        from .junctions import JunctionPolicy
        if not isinstance(policy, JunctionPolicy):
            raise TypeError("policy must be a JunctionPolicy")
        policy.validate()
        self._push(("capture_junctions", policy.replace()), f"capture junctions using policy {_rust_debug(policy.name)}")

    def relax_and_capture(self, *, iterations: int, capture_every: int, policy) -> None:
        """Relaxes for `iterations`, capturing junctions every `capture_every`
        (tangle_python/src/recipe.rs:573-598; tangle_generate/src/recipe.rs:1606-1651)."""
        from .junctions import JunctionPolicy
        iterations, capture_every = int(iterations), int(capture_every)
        if iterations <= 0 or capture_every <= 0 or capture_every > iterations:
            raise ValueError("iterations must be positive and capture_every must lie within 1..=iterations")
        if not isinstance(policy, JunctionPolicy):
            raise TypeError("policy must be a JunctionPolicy")
        policy.validate()
        self._push(("relax_and_capture", (iterations, capture_every, policy.replace())),
                   f"relax for {iterations} iterations and capture junctions every {capture_every}")

    def set_min_bend_radius(self, material, min_bend_radius: float) -> None:
        """Sets the minimum bend radius of every fiber of `material` (a
        Material or its name) for the rest of the run; the resident world
        takes the new curvature limits at once (SetMaterialBendRadius,
        tangle_python/src/recipe.rs:365-392; executed as in
        tangle_generate/src/recipe.rs:957-1026)."""
        if isinstance(material, Material):
            name = material.name
        elif isinstance(material, str):
            name = material
        else:
            raise TypeError("material must be a Material or a material name")
        if not name.strip():
            raise ValueError("material name must not be empty")
        radius = float(min_bend_radius)
        if not (radius > 0.0 and math.isfinite(radius)):
            raise ValueError("min_bend_radius must be positive and finite")
        self._push(("bend_radius", (name, radius)),
                   f"set {_rust_debug(name)} min bend radius to {_rust_display(radius)}")

    def fit_cell_to_active_fibers(self, *, axes=None, padding: float = 0.0) -> None:
        """Shrinks the cell along `axes` (default: the stack axis) to the
        active fibers' capsules plus `padding`, moving those walls in
        (FitCellToActiveFibers, tangle_python/src/recipe.rs:513-543; executed
        as in tangle_generate/src/recipe.rs:1206-1252 with
        recipe/compaction_control.rs:446-483)."""
        if axes is None:
            mask = [False, False, False]
            mask[self._stack_axis] = True
        else:
            mask = _axis_set(axes)
        if not any(mask):
            raise ValueError("at least one cell-fit axis must be selected")
        padding = float(padding)
        if not (padding >= 0.0 and math.isfinite(padding)):
            raise ValueError("padding must be nonnegative and finite")
        for k in range(3):
            if mask[k] and self._cell.periodic[k]:
                raise ValueError(f"cannot fit periodic cell axis {k} to active fibers")
        label = "".join("xyz"[k] for k in range(3) if mask[k])
        self._push(("fit_cell", (tuple(mask), padding)), f"fit cell axes {label} to active fibers")

    def operations(self) -> list[str]:
        return list(self._descriptions)

    def centerlines(self):
        return [f.placed.tolist() for f in self._fibers]

    def _packed(self, pin_fiber_ends: bool) -> _native.PackedAssembly:
        counts = np.array([f.placed.shape[0] for f in self._fibers], dtype=np.uint32)
        placed = np.concatenate([f.placed for f in self._fibers])
        intrinsic = np.concatenate([f.intrinsic for f in self._fibers])
        radius = np.array([f.radius for f in self._fibers])
        bend = np.array([f.material.min_bend_radius or 0.0 for f in self._fibers])
        if any(f.material.is_oval for f in self._fibers):
            raise ValueError("oval sections are not supported by the hip backend")
        steps = layers = None
        if any(f.step for f in self._fibers):
            steps = np.array([f.step for f in self._fibers], dtype=np.uint32)
        if any(f.formation_layer is not None for f in self._fibers):
            layers = np.array([0xFFFFFFFF if f.formation_layer is None else f.formation_layer
                               for f in self._fibers], dtype=np.uint32)
        return _native.PackedAssembly(counts, placed, intrinsic, radius, bend, self._cell.origin,
                                      self._cell.lengths, self._cell.periodic, pin_fiber_ends=pin_fiber_ends,
                                      formation_steps=steps, formation_layers=layers)

    def _fingerprint(self, packed: _native.PackedAssembly) -> str:
        """Identifies the fibers a checkpoint belongs to."""
        import hashlib
        h = hashlib.sha256()
        for arr in (packed.counts, packed.intrinsic, packed.placed, packed.radius,
                    packed.min_bend_radius if packed.min_bend_radius is not None else np.zeros(0),
                    packed.formation_steps if packed.formation_steps is not None else np.zeros(0),
                    packed.formation_layers if packed.formation_layers is not None else np.zeros(0)):
            h.update(np.ascontiguousarray(arr).tobytes())
        return h.hexdigest()

    def run(self, settings: RelaxationSettings | None = None, *, warm_start=None,
            checkpoint: "CheckpointSettings | None" = None) -> RunResult:
        """Runs every recorded operation on one resident device world, then
        relaxes on until the run's tolerances hold (recipe.rs:618-723,
        1123-1239; the controller tangle_generate/src/recipe.rs:857-1652 and
        the batch loop plugin.rs:1041-1181). With `checkpoint`, the world and
        the recipe's progress are saved every `interval_iterations`, at the
        end and on failure, and a run with `checkpoint.resume` continues the
        saved one (tangle_checkpoint/src/lib.rs:196-375); the recipe must be
        the one that wrote it. `warm_start` (a tangle_amd.learned.WarmStart)
        adds learned proposals, which have no counterpart in the reference."""
        from . import checkpoint as ckpt
        from . import formation
        if not self._fibers:
            raise ValueError("recipe contains no fibers")
        if not self._operations:
            raise ValueError("recipe contains no operations")
        if checkpoint is not None and not isinstance(checkpoint, ckpt.CheckpointSettings):
            raise TypeError("checkpoint must be CheckpointSettings")
        settings = settings or RelaxationSettings()
        started = time.perf_counter()
        d_min = min(2.0 * f.radius for f in self._fibers)
        config = settings.to_config(d_min)
        operations = list(self._operations)
        descriptions = list(self._descriptions)
        if all(op[0] == "activate" for op in operations):
            # a recipe that only inserts fibers relaxes them (recipe.rs:681-690)
            operations.append(("relax_until_converged", settings.max_iterations))
            descriptions.append(f"relax until converged (up to {settings.max_iterations} iterations)")
        axis = self._stack_axis
        initial_targets = None
        if any(op[0] in ("scale_layers", "place_layer") for op in operations):
            initial_targets = formation.layer_targets([f.placed for f in self._fibers],
                                                      [f.formation_layer for f in self._fibers], axis)
        packed = self._packed(settings.pin_fiber_ends)
        starts = np.concatenate([[0], np.cumsum(packed.counts.astype(np.int64))[:-1]])
        fingerprint = self._fingerprint(packed)
        run = _RunState(settings, config, warm_start)
        state = {"active_step": None, "targets": None if initial_targets is None else list(initial_targets),
                 # the host cell (fit_cell_to_active_fibers moves it) and the
                 # per-fiber bend radii (set_min_bend_radius changes them)
                 "cell": (list(self._cell.origin), list(self._cell.lengths)),
                 "bend": [f.material.min_bend_radius for f in self._fibers]}
        report = {"resumed": False, "resumed_iteration": None, "saves": 0, "last_iteration": None, "last_bytes": None}
        start_index, first_begin, restored = 0, None, None
        if checkpoint is not None and checkpoint.resume:
            path, case_id = checkpoint.source
            restored, progress, extra_devices = ckpt.load_with_extras(path, case_id)
            if progress.get("fingerprint") != fingerprint:
                raise ckpt.CheckpointError("the checkpoint was written by another recipe (its fibers differ)")
            run.total = restored.total_iterations
            run.last_saved = run.total
            report.update(resumed=True, resumed_iteration=run.total, last_iteration=run.total)
            if checkpoint.fresh_formation_on_resume:
                restored.layer_target = restored.vertex_target = None
                # the world restores onto the checkpoint's (possibly compacted) cell
                if restored.cell_lower is not None and all(np.isfinite(restored.cell_lower)):
                    state["cell"] = ([float(v) for v in restored.cell_lower],
                                     [float(np.float32(np.float32(restored.cell_upper[k])
                                                       - np.float32(restored.cell_lower[k]))) for k in range(3)])
            else:
                if progress.get("failure"):
                    raise RecipeError(progress["operation"], descriptions[progress["operation"]]
                                      if progress["operation"] < len(descriptions) else "run", run.total,
                                      progress["failure"])
                start_index, first_begin = int(progress["operation"]), int(progress["operation_begin"])
                run.events = list(progress["events"])
                run.warnings = list(progress["warnings"])
                run.needling = [tuple(n) for n in progress["needling"]]
                run.post_recipe_iterations = int(progress.get("post_recipe_iterations", 0))
                state["active_step"] = progress["active_step"]
                if progress.get("cell") is not None:
                    state["cell"] = (list(progress["cell"][0]), list(progress["cell"][1]))
                if progress.get("bend") is not None:
                    state["bend"] = list(progress["bend"])
                if progress.get("junctions"):
                    from .junctions import Junction
                    run.junctions = [Junction(int(i), law, int(ps), tuple((int(f), float(a)) for f, a in anchors))
                                     for i, law, ps, anchors in progress["junctions"]]
                    run.junction_captures = list(progress.get("junction_captures", []))
                if progress["layer_targets"] is not None:
                    state["targets"] = [np.float32(t) for t in progress["layer_targets"]]
                run.compactions = list(progress.get("compactions", []))
                run.compaction_steps = list(progress.get("compaction_steps", []))
                if progress.get("compaction_state") is not None:
                    # the compaction's loop continues from the batch status it last saw
                    run.resume_compaction = run.compaction_from_state(progress["compaction_state"],
                                                                      extra_devices.get("accepted"))
                    run.status = _status_from_json(progress["status"])
                    run.status_current = bool(progress.get("status_current", False))
                    run.converged = bool(progress.get("converged", False))
        world = (_native.World(packed, config) if restored is None
                 else _native.World.restore(packed, config, restored))
        holder = {"world": world}

        def restore_world(ck):
            # roll the resident world back to an accepted state (DeviceWorld::restore)
            holder["world"].close()
            holder["world"] = _native.World.restore(packed, config, ck)
            if state["bend"] != [f.material.min_bend_radius for f in self._fibers]:
                holder["world"].set_fiber_maximum_curvatures(_curvature_limits(state["bend"]))
            return holder["world"]
        if restored is not None:
            run.active_segments = int(((packed.counts.astype(np.int64) - 1) * (restored.fiber_active != 0)).sum())
            if state["bend"] != [f.material.min_bend_radius for f in self._fibers]:
                # curvature limits are recipe state, not device checkpoint state
                world.set_fiber_maximum_curvatures(_curvature_limits(state["bend"]))

        def save(failure=None):
            progress = {"fingerprint": fingerprint, "operation": run.op_index, "operation_begin": run.op_begin,
                        "events": run.events, "warnings": run.warnings, "needling": run.needling,
                        "active_step": state["active_step"], "post_recipe_iterations": run.post_recipe_iterations,
                        "layer_targets": None if state["targets"] is None else [float(t) for t in state["targets"]],
                        "cell": [list(state["cell"][0]), list(state["cell"][1])], "bend": list(state["bend"]),
                        "junctions": [[j.id, j.law, j.parameter_set, [list(a) for a in j.anchors]]
                                      for j in run.junctions],
                        "junction_captures": list(run.junction_captures),
                        "compactions": list(run.compactions), "compaction_steps": list(run.compaction_steps),
                        "status": None if run.status is None else _status_json(run.status),
                        "status_current": bool(run.status_current), "converged": bool(run.converged),
                        "failure": failure}
            compaction, accepted = run.compaction_state()
            extras = {}
            if compaction is not None:
                progress["compaction_state"] = compaction
                if accepted is not None:
                    extras["accepted"] = accepted
            nbytes = ckpt.save(checkpoint.path, checkpoint.case_id, holder["world"].checkpoint(), progress,
                               extras=extras)
            report.update(saves=report["saves"] + 1, last_iteration=run.total, last_bytes=nbytes)

        if checkpoint is not None:
            run.checkpoint_interval = int(checkpoint.interval_iterations)
            run.checkpoint_hook = save
        try:
            info = world.info()
            run.uploaded = int(info.device_bytes)
            ctx = {"state": state, "fibers": self._fibers, "stack_axis": axis, "periodic": self._cell.periodic,
                   "restore": restore_world, "holder": holder, "counts": packed.counts,
                   "intrinsic": np.asarray(packed.intrinsic, dtype=np.float32).reshape(-1, 3)}
            try:
                for index in range(start_index, len(operations)):
                    kind, arg = operations[index]
                    begin = first_begin if index == start_index and first_begin is not None else run.total
                    run.op_index, run.op_begin = index, begin
                    ctx["world"] = holder["world"]
                    self._run_operation(holder["world"], run, state, index, descriptions[index], kind, arg, begin,
                                        initial_targets, starts, ctx)
                resumed_in_finish = start_index == len(operations) and first_begin is not None
                run.op_index, run.op_begin = len(operations), first_begin if resumed_in_finish else run.total
                run.finish(holder["world"], begin=run.op_begin)
            except RecipeError as err:
                run.save_checkpoint(failure=err.reason)
                raise
            run.save_checkpoint()
            world = holder["world"]
            info = world.info()
            positions = world.download_positions()
            run.downloaded += positions.nbytes
            rebuilds = int(world.info().neighbor_list_rebuilds)
        finally:
            holder["world"].close()
        status = run.status
        wall = time.perf_counter() - started
        return RunResult(
            fiber_count=len(self._fibers), iterations=run.total, converged=run.converged,
            max_penetration=float(status.max_penetration) if status else float("inf"),
            max_curvature_ratio=float(status.max_curvature_ratio) if status else float("inf"),
            max_displacement=float(status.max_displacement) if status else 0.0,
            active_segments=int(info.segment_count), active_vertices=int(info.vertex_count),
            cell_count=int(info.cell_count), uploaded_bytes=run.uploaded, downloaded_bytes=run.downloaded,
            events=run.events + ([f"warm start: {run.proposals_kept} proposals kept, "
                                  f"{run.proposals_reverted} reverted"] if warm_start else []),
            warnings=run.warnings, neighbor_list_rebuilds=rebuilds, wall_seconds=wall,
            positions=positions, vertex_counts=packed.counts.copy(),
            post_recipe_iterations=run.post_recipe_iterations, segment_iterations=run.segment_iterations,
            needling=list(run.needling), compactions=list(run.compactions), junctions=list(run.junctions),
            junction_captures=list(run.junction_captures),
            compaction_steps=list(run.compaction_steps), resumed=report["resumed"], resumed_iteration=report["resumed_iteration"],
            checkpoint_saves=report["saves"], last_checkpoint_iteration=report["last_iteration"],
            last_checkpoint_bytes=report["last_bytes"], _fibers=list(self._fibers),
            _cell=Cell(state["cell"][1], "".join("xyz"[k] for k in range(3) if self._cell.periodic[k]),
                       state["cell"][0], stack_axis=self._cell.stack_axis),
            _bend=list(state["bend"]),
        )

    def _run_operation(self, world, run, state, index, desc, kind, arg, begin, initial_targets, starts, ctx=None):
        """One recorded operation (control_formation_recipe,
        tangle_generate/src/recipe.rs:857-1652)."""
        from . import formation
        axis = self._stack_axis
        if kind not in ("solve", "relax_for", "relax_until_converged", "settle_targets"):
            run.status_current = False  # the operation may move or (de)activate material
        if kind == "activate":
            segments, vertices = world.activate_formation_step(arg)
            run.active_segments = segments
            state["active_step"] = arg
            run.converged = False
            run.events.append(f"insert fiber groups through step {arg} "
                              f"({segments} segments, {vertices} vertices active)")
        elif kind == "solve":
            run.solve(world, index, desc, *arg, begin=begin)
        elif kind in ("relax_for", "relax_until_converged"):
            run.relax(world, index, desc, kind, arg, begin=begin)
        elif kind == "settle_targets":
            run.settle(world, index, desc, *arg, begin=begin)
        elif kind == "scale_layers":
            factor, stiffness, max_translation = arg
            origin, lengths = state["cell"]
            center = np.float32(origin[axis] + 0.5 * lengths[axis])
            state["targets"] = [np.float32(center + np.float32(np.float32(t - center) * np.float32(factor)))
                                for t in initial_targets]
            world.set_layer_targets(axis, state["targets"], 0, len(state["targets"]) - 1, stiffness,
                                    max_translation)
            run.converged = False
            run.events.append(f"move layers to spacing scale {factor:.3f}")
        elif kind == "place_layer":
            layer, gap, stiffness, max_translation = arg
            targets = state["targets"]
            if layer >= len(targets):
                raise RecipeError(index, desc, run.total,
                                  f"cannot place layer {layer}: only {len(targets)} layers exist")
            targets[layer] = np.float32(targets[layer - 1] + np.float32(gap))
            world.set_layer_targets(axis, targets, layer, layer, stiffness, max_translation)
            run.converged = False
            run.events.append(f"place layer {layer} at gap {gap:.3e} above layer {layer - 1}")
        elif kind == "release_layers":
            world.clear_layer_targets()
            run.converged = False
            run.events.append("release manufacturing-layer targets")
        elif kind == "needle":
            n = arg
            active_step = state["active_step"]
            if active_step is None or active_step < n.layer:
                raise RecipeError(index, desc, run.total,
                                  f"cannot needle layer {n.layer}: it has not been inserted yet")
            positions = world.download_positions().reshape(-1, 3)
            run.downloaded += positions.nbytes
            origin, lengths = state["cell"]
            lower = [np.float32(o) for o in origin]
            upper = [np.float32(o + l) for o, l in zip(origin, lengths)]
            vertices = formation.select_needle_vertices(self._fibers, positions, starts, active_step, axis,
                                                        self._cell.periodic, lower, upper, n)
            world.clear_vertex_targets()
            if vertices:
                owner = np.searchsorted(starts, np.asarray(vertices), side="right") - 1
                diameter = min(2.0 * self._fibers[int(f)].radius for f in owner)
                limit = np.float32(np.float32(diameter) * np.float32(n.max_translation_over_diameter))
                world.set_vertex_targets(axis, vertices, -n.depth, n.stiffness,
                                         min(np.float32(n.max_translation), limit))
            run.converged = False
            run.needling.append((n.layer, len(vertices), n.depth))
            run.events.append(f"needle {len(vertices)} vertices in layer {n.layer} through depth "
                              f"{n.depth:.3e}")
        elif kind == "release_needles":
            world.clear_vertex_targets()
            run.converged = False
            run.events.append("release needle targets")
        elif kind == "fit_cell":
            mask, padding = arg
            positions = world.download_positions().reshape(-1, 3)
            run.downloaded += positions.nbytes
            active_step = state["active_step"]
            active = [active_step is None or f.step <= active_step for f in self._fibers]
            fit_lower, fit_upper = _active_capsule_bounds(self._fibers, positions, starts, active, padding)
            if fit_lower is None:
                raise RecipeError(index, desc, run.total, "cannot fit a cell without active fiber segments")
            old_lower, old_upper = world.cell_bounds()
            new_lower = [np.float32(max(fit_lower[k], np.float32(old_lower[k]))) if mask[k]
                         else np.float32(old_lower[k]) for k in range(3)]
            new_upper = [np.float32(min(fit_upper[k], np.float32(old_upper[k]))) if mask[k]
                         else np.float32(old_upper[k]) for k in range(3)]
            try:
                world.compact_cell(new_lower, new_upper, "moving_walls")
            except (ValueError, _native.NativeError) as err:
                raise RecipeError(index, desc, run.total, f"cannot fit the cell: {err}") from None
            # update_host_cell (compaction_control.rs:476-483)
            extent = [np.float32(new_upper[k] - new_lower[k]) for k in range(3)]
            state["cell"] = ([float(v) for v in new_lower], [float(v) for v in extent])
            run.converged = False
            from .export import rust_exp
            run.events.append(f"fit cell to active fibers with {rust_exp(np.float32(padding), 3)} padding: "
                              f"[{', '.join(rust_exp(e, 3) for e in extent)}]")
        elif kind == "capture_junctions":
            report = run.capture_junctions(world, index, desc, arg, ctx)
            run.events.append(f"capture {report['created']} junctions with '{report['policy_name']}' from "
                              f"{report['candidates']} candidates")
        elif kind == "relax_and_capture":
            iterations, every, policy = arg
            config = _native.Config.from_buffer_copy(run.config)
            config.force_full_iterations = 1
            last_capture = None
            while True:
                completed = run.total - begin
                if completed > 0 and completed % every == 0 and last_capture != run.total:
                    report = run.capture_junctions(world, index, desc, policy, ctx)
                    run.events.append(f"periodic capture created {report['created']} '{report['policy_name']}' "
                                      f"junctions from {report['candidates']} candidates")
                    last_capture = run.total
                if completed >= iterations:
                    run.events.append(f"relax for {iterations} iterations with junction capture every {every}")
                    break
                next_capture = (completed // every + 1) * every
                remaining = run._budget(index, desc)
                batch = max(1, min(remaining, run.settings.iterations_per_batch,
                                   min(next_capture, iterations) - completed))
                run.converged = False
                run._batch(world, config, batch)
        elif kind == "compact":
            spec, overrides = arg
            run.compact(ctx, index, desc, spec, overrides)
        elif kind == "bend_radius":
            name, radius = arg
            matching = [i for i, f in enumerate(self._fibers) if f.material.name == name]
            if not matching:
                names = list(dict.fromkeys(f.material.name for f in self._fibers))
                reason = (f"material {_rust_debug(name)} has no fibers" if name in names else
                          f"unknown material {_rust_debug(name)}; known materials: "
                          f"{', '.join(_rust_debug(n) for n in names)}")
                raise RecipeError(index, desc, run.total, reason)
            for i in matching:
                state["bend"][i] = radius
            world.set_fiber_maximum_curvatures(_curvature_limits(state["bend"]))
            run.converged = False
            from .export import rust_exp
            run.events.append(f"set {len(matching)} {_rust_debug(name)} fibers to minimum bend radius "
                              f"{rust_exp(radius, 3)}")
        else:
            raise AssertionError(kind)


class _BatchedProposals:
    """Learned proposals on a batched world: the single-world rule of
    _RunState._propose applied per specimen (no counterpart in the
    reference)."""

    def __init__(self, warm_start, config):
        self.ws = warm_start
        self.config = config
        self.marked = False
        self.last = 0
        self.kept = self.reverted = 0

    def _residual(self, st):
        pen_tol = max(self.config.penetration_tolerance, 1e-30)
        curv_tol = max(self.config.curvature_ratio_tolerance, 1e-30)
        return max(st.max_penetration / pen_tol, (st.max_curvature_ratio - 1.0) / curv_tol)

    def step(self, world, done, batch, room):
        """Proposes when one is due and its settle iterations leave one of
        the `room` the budget has left; returns the iteration count after
        the proposal and the next batch size, cut so that the following
        proposal falls on a batch boundary and the budget holds."""
        from . import learned
        ws = self.ws
        if done >= ws.active_iterations:
            return done, batch
        if not self.marked:
            learned.mark_history(world)
            self.marked, self.last = True, done
        elif done - self.last >= ws.interval and ws.settle < room:
            before = world.specimen_status()
            learned.propose(world, ws.model, ws.scale)
            # specimens already converged keep their positions
            learned.revert_specimens(world, [bool(st.converged) for st in before])
            if ws.settle > 0:
                world.run_batch(self.config, ws.settle)
                after = world.specimen_status()
            else:
                after = before
            worse = [not b.converged and self._residual(a) > ws.tolerance * self._residual(b)
                     for b, a in zip(before, after)]
            if any(worse):
                learned.revert_specimens(world, worse)
            self.reverted += sum(worse)
            self.kept += sum(1 for b, w in zip(before, worse) if not b.converged and not w)
            learned.mark_history(world)
            self.last = int(max(st.total_iterations for st in world.specimen_status()))
            room -= self.last - done
            done = self.last
        elif done - self.last >= ws.interval:
            return done, batch
        return done, max(1, min(batch, room, self.last + ws.interval - done))


def relax_batch(recipes, settings: RelaxationSettings | None = None, *, max_iterations: int | None = None,
                iterations: int | None = None, autotune: bool = False, warm_start=None) -> list[RunResult]:
    """Relaxes independent specimens together (no counterpart in the
    reference, which runs one world at a time): every recipe's fibers and
    cell become one specimen of a batched device world, all specimens step
    through the same launches, and each stops on its own convergence gate.

    With `iterations`, every specimen runs exactly that many iterations
    (force_full_iterations); otherwise each relaxes until converged, within
    `max_iterations` (default: settings.max_iterations). With `autotune`,
    the cell size, skin and launch shapes are first picked by
    tangle_amd.autotune on a scratch copy (same positions bit for bit,
    fewer seconds). With `warm_start` (a tangle_amd.learned.WarmStart) every
    specimen still relaxing takes the network's proposal every `interval`
    iterations, runs `settle` iterations and keeps it only when its own
    normalised residual did not rise; converged specimens are left as they
    are. Returns one RunResult per recipe, in order; a specimen that does not
    converge in the budget is returned with converged=False rather than
    raising."""
    recipes = list(recipes)
    if not recipes:
        raise ValueError("relax_batch needs at least one recipe")
    for r in recipes:
        if not isinstance(r, Recipe):
            raise TypeError("relax_batch takes Recipe objects")
        if not r._fibers:
            raise ValueError("a recipe in the batch contains no fibers")
    settings = settings or RelaxationSettings()
    budget = int(max_iterations if max_iterations is not None else settings.max_iterations)
    if budget <= 0 or (iterations is not None and iterations <= 0):
        raise ValueError("iteration counts must be positive")
    started = time.perf_counter()
    d_min = min(2.0 * f.radius for r in recipes for f in r._fibers)
    config = settings.to_config(d_min)
    if iterations is not None:
        config.force_full_iterations = 1
        budget = int(iterations)
    packed = [r._packed(settings.pin_fiber_ends) for r in recipes]
    tuning = None
    if autotune:
        from .autotune import autotune as _autotune
        tuning = _autotune(packed, config)
        config = tuning.configure(config)
    with _native.World.batch(packed, config) as world:
        if tuning is not None:
            tuning.apply(world)
        info = world.info()
        done = 0
        proposals = _BatchedProposals(warm_start, config) if warm_start is not None and iterations is None else None
        while done < budget:
            batch = min(settings.iterations_per_batch, budget - done)
            if proposals is not None:
                done, batch = proposals.step(world, done, batch, budget - done)
            status = world.run_batch(config, batch)
            done = int(status.total_iterations)
            if status.converged or status.batch_iterations == 0:
                break
        statuses = world.specimen_status()
        positions = world.specimen_positions()
        rebuilds = int(world.info().neighbor_list_rebuilds)
    wall = time.perf_counter() - started
    out = []
    for r, p, st, pos in zip(recipes, packed, statuses, positions):
        out.append(RunResult(
            fiber_count=len(r._fibers), iterations=int(st.total_iterations), converged=bool(st.converged),
            max_penetration=float(st.max_penetration), max_curvature_ratio=float(st.max_curvature_ratio),
            max_displacement=float(st.max_displacement), active_segments=int(p.vertex_count - p.counts.size),
            active_vertices=int(p.vertex_count), cell_count=int(info.cell_count), uploaded_bytes=0,
            downloaded_bytes=int(pos.nbytes), events=[f"batched with {len(recipes)} specimens"], warnings=[],
            neighbor_list_rebuilds=rebuilds, wall_seconds=wall, positions=pos, vertex_counts=p.counts.copy(),
            _fibers=list(r._fibers), _cell=r._cell))
    return out
