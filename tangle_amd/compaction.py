"""Closed-loop cell compaction: the settings of ``Recipe.compact`` and the
controller that runs it between device batches.

The controller follows the reference's compact-relax-measure loop
(tangle_generate/src/recipe/compaction_control.rs:5-568, with the targets,
paths and step rules of tangle_generate/src/compaction.rs and the Python
settings of tangle_python/src/compaction.rs): relax to an admissible
baseline, then repeatedly shorten the cell by an adaptive logarithmic strain
split among the axes by the path rule, let the material follow
(`World.compact_cell`), relax for a window, and accept the step when the
window converges or roll the world back to the last accepted state and halve
the step when it does not. Pressures and energies come from
`World.compaction_metrics`. Scalars keep the reference's precision: cell
lengths and volumes in f64, strains, pressures and work in f32.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace as _replace

import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)
KINEMATICS = ("rigid_fiber_centers", "moving_walls", "affine_vertices")


def _f32(x) -> np.float32:
    return np.float32(x)


def _axes(value, name: str) -> tuple:
    """An axis mask from "xz", (True, False, True), an axis letter or index
    (parse_axis_mask, tangle_python/src/common.rs:43-80)."""
    if isinstance(value, str):
        mask = [False, False, False]
        for letter in value.lower():
            if letter not in "xyz":
                raise ValueError(f"{name} letters must be drawn from 'xyz'; got {value!r}")
            mask["xyz".index(letter)] = True
        return tuple(mask)
    if isinstance(value, bool):
        return (value,) * 3
    if isinstance(value, int):
        if value not in (0, 1, 2):
            raise ValueError(f"{name} axis must be 0, 1 or 2")
        return tuple(k == value for k in range(3))
    mask = tuple(bool(v) for v in value)
    if len(mask) != 3:
        raise ValueError(f"{name} must name three axes")
    return mask


def _positive_finite(value, name):
    if not (math.isfinite(value) and value > 0.0):
        raise ValueError(f"{name} must be positive and finite")


def _nonnegative_active(values, name):
    values = tuple(float(v) for v in values)
    if len(values) != 3 or any(not math.isfinite(v) or v < 0.0 for v in values) or not any(v > 0.0 for v in values):
        raise ValueError(f"{name} must be nonnegative with at least one positive axis")
    return values


def _require_active(axes):
    if not any(axes):
        raise ValueError("at least one compaction axis must be active")


def _display(value: float) -> str:
    return np.format_float_positional(float(value), unique=True, trim="-")


# ---- targets (compaction.rs:7-21; tangle_python/src/compaction.rs:24-160) ----

@dataclass(frozen=True)
class VolumeFractionTarget:
    """Stop when nominal fiber volume over cell volume reaches `value` (< 1)."""

    value: float

    def __post_init__(self):
        _positive_finite(self.value, "value")
        if self.value >= 1.0:
            raise ValueError("volume fraction target must be less than 1")

    def __repr__(self):
        return f"VolumeFractionTarget({_display(self.value)})"


@dataclass(frozen=True)
class CellVolumeTarget:
    """Stop when the cell volume reaches `value`."""

    value: float

    def __post_init__(self):
        _positive_finite(self.value, "value")

    def __repr__(self):
        return f"CellVolumeTarget({_display(self.value)})"


@dataclass(frozen=True)
class MeanPressureTarget:
    """Stop when the mean of the three directional penalty pressures reaches `value`."""

    value: float

    def __post_init__(self):
        _positive_finite(self.value, "value")

    def __repr__(self):
        return f"MeanPressureTarget({_display(self.value)})"


@dataclass(frozen=True)
class PenaltyEnergyTarget:
    """Stop when total contact, stretch and excess-bending energy reaches `value`."""

    value: float

    def __post_init__(self):
        _positive_finite(self.value, "value")

    def __repr__(self):
        return f"PenaltyEnergyTarget({_display(self.value)})"


@dataclass(frozen=True)
class CellLengthsTarget:
    """Stop when every cell edge is at most its target length."""

    lengths: tuple

    def __post_init__(self):
        lengths = tuple(float(v) for v in self.lengths)
        if len(lengths) != 3:
            raise ValueError("lengths must hold three values")
        for v in lengths:
            _positive_finite(v, "lengths")
        object.__setattr__(self, "lengths", lengths)

    def __repr__(self):
        return f"CellLengthsTarget(lengths=[{', '.join(_display(v) for v in self.lengths)}])"


@dataclass(frozen=True)
class DirectionalPressureTarget:
    """Stop when every axis with a nonzero target carries that pressure."""

    pressures: tuple

    def __post_init__(self):
        object.__setattr__(self, "pressures", _nonnegative_active(self.pressures, "pressures"))

    def __repr__(self):
        return f"DirectionalPressureTarget(pressures=[{', '.join(_display(v) for v in self.pressures)}])"


TARGETS = (VolumeFractionTarget, CellVolumeTarget, CellLengthsTarget, MeanPressureTarget,
           DirectionalPressureTarget, PenaltyEnergyTarget)


# ---- paths (compaction.rs:23-57; tangle_python/src/compaction.rs:196-330) ----

@dataclass(frozen=True)
class AxisWeightsPath:
    """Fixed nonnegative logarithmic-strain ratios; None: the stack axis."""

    weights: tuple | None = None

    def __init__(self, weights=None):
        if weights is not None:
            if isinstance(weights, str) or isinstance(weights, bool) or isinstance(weights, int):
                weights = tuple(1.0 if a else 0.0 for a in _axes(weights, "weights"))
            weights = _nonnegative_active(weights, "weights")
        object.__setattr__(self, "weights", weights)


@dataclass(frozen=True)
class EqualPressurePath:
    """Adapt strain rates toward equal pressure on the selected axes."""

    axes: tuple
    pressure_floor: float = 1.0e-12

    def __init__(self, axes, *, pressure_floor: float = 1.0e-12):
        axes = _axes(axes, "axes")
        _require_active(axes)
        _positive_finite(pressure_floor, "pressure_floor")
        object.__setattr__(self, "axes", axes)
        object.__setattr__(self, "pressure_floor", float(pressure_floor))


@dataclass(frozen=True)
class StressRatioPath:
    """Adapt strain rates toward the given directional pressure ratio."""

    ratio: tuple
    pressure_floor: float = 1.0e-12

    def __init__(self, ratio, *, pressure_floor: float = 1.0e-12):
        ratio = _nonnegative_active(ratio, "ratio")
        _positive_finite(pressure_floor, "pressure_floor")
        object.__setattr__(self, "ratio", ratio)
        object.__setattr__(self, "pressure_floor", float(pressure_floor))


@dataclass(frozen=True)
class MinimumWorkPath:
    """Put each increment on the currently least loaded selected axis."""

    axes: tuple

    def __init__(self, axes):
        axes = _axes(axes, "axes")
        _require_active(axes)
        object.__setattr__(self, "axes", axes)


PATHS = (AxisWeightsPath, EqualPressurePath, StressRatioPath, MinimumWorkPath)


# ---- settings (tangle_python/src/compaction.rs:407-700; defaults compaction.rs:59-165) ----

@dataclass
class CompactionSettings:
    """One closed-loop compaction command (CompactionSettings). Defaults:
    VolumeFractionTarget(0.3), the stack axis, rigid fiber centres, both
    faces moving equally."""

    target: object = field(default_factory=lambda: VolumeFractionTarget(0.3))
    path: object = field(default_factory=AxisWeightsPath)
    kinematics: str = "rigid_fiber_centers"
    cell_anchor: tuple = (0.5, 0.5, 0.5)
    balance_opposing_faces: bool = False
    face_pressure_floor: float = 1.0e-12
    face_balance_strength: float = 0.5
    initial_log_strain: float = 0.02
    min_log_strain: float = 0.001
    max_log_strain: float = 0.05
    growth_factor: float = 1.25
    shrink_factor: float = 0.5
    relax_iterations: int = 64
    max_shortening_over_min_diameter: float = 0.5
    max_penetration: float = 2.0e-4
    max_curvature_ratio: float = 1.001
    max_pressure: float = math.inf
    max_penalty_energy: float = math.inf
    max_steps: int = 1000
    max_relax_windows: int = 4
    contact_energy_stiffness: float = 1.0
    stretch_energy_stiffness: float = 1.0
    bending_energy_stiffness: float = 1.0
    target_tolerance: float = 1.0e-3

    def __post_init__(self):
        self.validate()

    @classmethod
    def volume_fraction(cls, target: float, **changes) -> "CompactionSettings":
        return cls(VolumeFractionTarget(float(target)), **changes)

    def replace(self, **changes) -> "CompactionSettings":
        return _replace(self, **changes)

    def copy(self) -> "CompactionSettings":
        return _replace(self)

    def validate(self) -> None:
        if not isinstance(self.target, TARGETS):
            raise TypeError("target must be a VolumeFractionTarget, CellVolumeTarget, CellLengthsTarget, "
                            "MeanPressureTarget, DirectionalPressureTarget or PenaltyEnergyTarget")
        if not isinstance(self.path, PATHS):
            raise TypeError("path must be an AxisWeightsPath, EqualPressurePath, StressRatioPath, "
                            "or MinimumWorkPath")
        if self.kinematics not in KINEMATICS:
            raise ValueError(f"kinematics must be one of {', '.join(KINEMATICS)}")
        anchor = tuple(float(v) for v in self.cell_anchor)
        if len(anchor) != 3 or any(not math.isfinite(v) or not 0.0 <= v <= 1.0 for v in anchor):
            raise ValueError("cell_anchor values must be in [0, 1]")
        self.cell_anchor = anchor
        _positive_finite(self.face_pressure_floor, "face_pressure_floor")
        if not (math.isfinite(self.face_balance_strength) and 0.0 <= self.face_balance_strength <= 1.0):
            raise ValueError("face_balance_strength must be in [0, 1]")
        for name in ("initial_log_strain", "min_log_strain", "max_log_strain", "growth_factor", "shrink_factor",
                     "max_shortening_over_min_diameter", "max_penetration", "max_curvature_ratio",
                     "contact_energy_stiffness", "stretch_energy_stiffness", "bending_energy_stiffness",
                     "target_tolerance"):
            _positive_finite(float(getattr(self, name)), name)
        if self.growth_factor < 1.0:
            raise ValueError("growth_factor must be at least 1")
        if self.shrink_factor >= 1.0:
            raise ValueError("shrink_factor must be in (0, 1)")
        if self.max_curvature_ratio < 1.0:
            raise ValueError("max_curvature_ratio must be at least 1")
        if self.target_tolerance >= 1.0:
            raise ValueError("target_tolerance must be below 1")
        if not (self.min_log_strain <= self.initial_log_strain <= self.max_log_strain):
            raise ValueError("log strain increments must satisfy min <= initial <= max")
        if self.relax_iterations <= 0 or self.max_steps <= 0 or self.max_relax_windows <= 0:
            raise ValueError("compaction iteration and step counts must be positive")
        for name in ("max_pressure", "max_penalty_energy"):
            value = float(getattr(self, name))
            if math.isnan(value) or value <= 0.0:
                raise ValueError(f"{name} must be positive; infinity is allowed")

    def weights_for(self, stack_axis: int):
        """The path with its default resolved (Path::to_rust)."""
        if isinstance(self.path, AxisWeightsPath) and self.path.weights is None:
            return AxisWeightsPath(tuple(1.0 if k == stack_axis else 0.0 for k in range(3)))
        return self.path


# ---- the step rules (compaction.rs:199-378) ---------------------------------

def nominal_fiber_volume(fibers, active_through) -> float:
    """Cross-section area times intrinsic arc length, summed over the fibers
    formed through `active_through` (None: all), in f64 (compaction.rs:233-260)."""
    total = 0.0
    for f in fibers:
        if active_through is not None and f.step > active_through:
            continue
        pts = np.asarray(f.intrinsic, dtype=np.float64)
        length = 0.0
        for a, b in zip(pts[:-1], pts[1:]):
            d = b - a
            length += math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        total += math.pi * f.radius * f.radius * length
    return total


def axis_weights(path, pressure) -> list:
    """Normalised f32 strain weights of a path (axis_weights, compaction.rs:330-362)."""
    p = [_f32(v) for v in pressure]
    if isinstance(path, AxisWeightsPath):
        raw = [_f32(w) for w in path.weights]
    elif isinstance(path, EqualPressurePath):
        floor = _f32(path.pressure_floor)
        raw = [_f32(_f32(1.0) / _f32(p[k] + floor)) if path.axes[k] else _f32(0.0) for k in range(3)]
    elif isinstance(path, StressRatioPath):
        floor = _f32(path.pressure_floor)
        raw = [_f32(_f32(path.ratio[k]) / _f32(p[k] + floor)) if path.ratio[k] > 0.0 else _f32(0.0)
               for k in range(3)]
    else:
        active = [k for k in range(3) if path.axes[k]]
        selected = min(active, key=lambda k: _total_key(p[k]))  # the first of equal minima
        raw = [_f32(1.0 if k == selected else 0.0) for k in range(3)]
    total = _f32(_f32(raw[0] + raw[1]) + raw[2])
    return [_f32(w / total) for w in raw]


def _total_key(x: np.float32) -> int:
    i = int(np.float32(x).view(np.int32))
    return i ^ ((i >> 31) & 0x7FFFFFFF)


def diameter_limited_log_strain(requested, weights, lengths, minimum_diameter, ratio) -> np.float32:
    """Caps the strain so no weighted axis shortens by more than `ratio`
    minimum diameters (diameter_limited_log_strain, compaction.rs:364-380)."""
    maximum_shortening = minimum_diameter * float(_f32(ratio))
    limited = _f32(requested)
    for k in range(3):
        if weights[k] > 0.0:
            retained = min(max(1.0 - maximum_shortening / lengths[k], 2.2250738585072014e-308), 1.0)
            limited = _f32(min(limited, _f32(-math.log(retained) / float(weights[k]))))
    return _f32(max(limited, _f32(F32_EPS)))


def target_reached(target, tolerance, lengths, volume_fraction, pressure, total_energy) -> bool:
    """target_reached (compaction_control.rs:484-510)."""
    tol = float(_f32(tolerance))
    tol32 = _f32(tol)
    if isinstance(target, VolumeFractionTarget):
        return volume_fraction >= target.value * (1.0 - tol)
    if isinstance(target, CellVolumeTarget):
        return lengths[0] * lengths[1] * lengths[2] <= target.value * (1.0 + tol)
    if isinstance(target, CellLengthsTarget):
        return all(lengths[k] <= target.lengths[k] * (1.0 + tol) for k in range(3))
    one_minus = _f32(_f32(1.0) - tol32)
    if isinstance(target, MeanPressureTarget):
        mean = _f32(_f32(_f32(_f32(pressure[0]) + _f32(pressure[1])) + _f32(pressure[2])) / _f32(3.0))
        return mean >= _f32(_f32(target.value) * one_minus)
    if isinstance(target, DirectionalPressureTarget):
        return all(_f32(target.pressures[k]) == 0.0 or _f32(pressure[k]) >= _f32(_f32(target.pressures[k]) * one_minus)
                   for k in range(3))
    return _f32(total_energy) >= _f32(_f32(target.value) * one_minus)


def strain_direction_and_remaining(target, path, lengths, volume_fraction, pressure):
    """(weights, remaining log strain or None) (compaction_control.rs:512-540)."""
    if isinstance(target, CellLengthsTarget):
        remaining = [_f32(max(math.log(lengths[k] / target.lengths[k]), 0.0)) for k in range(3)]
        total = _f32(_f32(remaining[0] + remaining[1]) + remaining[2])
        return [_f32(r / total) for r in remaining], total
    weights = axis_weights(path, pressure)
    if isinstance(target, VolumeFractionTarget):
        return weights, _f32(max(math.log(target.value / volume_fraction), 0.0))
    if isinstance(target, CellVolumeTarget):
        return weights, _f32(max(math.log(lengths[0] * lengths[1] * lengths[2] / target.value), 0.0))
    return weights, None


def guard_reached(settings: CompactionSettings, max_penetration, max_curvature_ratio, pressure, total_energy,
                  steps) -> str | None:
    """guard_reached (compaction_control.rs:542-568)."""
    if steps >= settings.max_steps:
        return "maximum compaction steps"
    if _f32(max_penetration) > _f32(settings.max_penetration):
        return "maximum penetration"
    if _f32(max_curvature_ratio) > _f32(settings.max_curvature_ratio):
        return "max curvature ratio"
    if any(_f32(p) > _f32(settings.max_pressure) for p in pressure):
        return "maximum pressure"
    if _f32(total_energy) > _f32(settings.max_penalty_energy):
        return "maximum penalty energy"
    return None


def reason_text(reason) -> str:
    """CompactionStopReason as Rust's `{:?}` writes it."""
    if reason in ("TargetReached", "Jammed"):
        return reason
    return 'GuardReached("' + reason[1].replace("\\", "\\\\").replace('"', '\\"') + '")'


def rust_exp3(value) -> str:
    from .export import rust_exp
    return rust_exp(float(value), 3)
