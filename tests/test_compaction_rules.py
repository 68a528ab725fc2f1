"""The compaction controller's host rules (tangle_amd/compaction.py) against
the reference's own unit tests (tangle_generate/src/compaction.rs:383-432)
and closed forms; settings validation as the Python binding validates
(tangle_python/src/compaction.rs:629-700)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import compaction as cp


class Fiber:
    def __init__(self, intrinsic, radius, step=0):
        self.intrinsic = np.asarray(intrinsic, dtype=np.float64)
        self.radius = radius
        self.step = step


def test_nominal_volume_uses_intrinsic_arc_length_and_section_area():
    # compaction.rs:388-404: radius 0.5, intrinsic length 2 -> 0.5 pi
    f = Fiber([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]], 0.5)
    assert abs(cp.nominal_fiber_volume([f], None) - 0.5 * math.pi) < 1e-12
    late = Fiber([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]], 0.5, step=2)
    assert abs(cp.nominal_fiber_volume([f, late], 1) - 0.5 * math.pi) < 1e-12
    assert abs(cp.nominal_fiber_volume([f, late], None) - 1.25 * math.pi) < 1e-12


def test_minimum_work_selects_the_softest_active_axis():
    # compaction.rs:406-418
    w = cp.axis_weights(cp.MinimumWorkPath((True, False, True)), [2.0, 0.0, 1.0])
    assert [float(v) for v in w] == [0.0, 0.0, 1.0]
    # ties go to the first axis
    assert [float(v) for v in cp.axis_weights(cp.MinimumWorkPath("xyz"), [1.0, 1.0, 1.0])] == [1.0, 0.0, 0.0]


def test_diameter_limit_caps_cell_shortening():
    # compaction.rs:420-431
    strain = cp.diameter_limited_log_strain(0.01, [0.0, 0.0, 1.0], [1.0e-3] * 3, 7.0e-6, 0.25)
    shortening = 1.0e-3 * (1.0 - math.exp(-float(strain)))
    assert abs(shortening - 1.75e-6) < 1.0e-10
    # an unconstrained request passes through
    assert cp.diameter_limited_log_strain(2e-4, [1.0, 0.0, 0.0], [1.0] * 3, 1e-3, 0.5) == np.float32(2e-4)
    assert cp.diameter_limited_log_strain(1e-9, [1.0, 0.0, 0.0], [1.0] * 3, 1e-3, 0.5) == np.finfo(np.float32).eps


def test_path_weights_closed_forms():
    w = cp.axis_weights(cp.EqualPressurePath("xz", pressure_floor=1.0), [1.0, 5.0, 3.0])
    # 1/(1+1) : 0 : 1/(3+1) normalised
    assert np.allclose([float(v) for v in w], [2 / 3, 0.0, 1 / 3], rtol=1e-6)
    w = cp.axis_weights(cp.StressRatioPath((2.0, 1.0, 0.0), pressure_floor=1.0), [1.0, 1.0, 7.0])
    assert np.allclose([float(v) for v in w], [2 / 3, 1 / 3, 0.0], rtol=1e-6)
    w = cp.axis_weights(cp.AxisWeightsPath((1.0, 1.0, 2.0)), [0.0] * 3)
    assert [float(v) for v in w] == [0.25, 0.25, 0.5]


def test_targets_and_remaining_strain():
    lengths = [2.0, 2.0, 2.0]
    t = cp.VolumeFractionTarget(0.2)
    assert not cp.target_reached(t, 1e-3, lengths, 0.1, [0.0] * 3, 0.0)
    assert cp.target_reached(t, 1e-3, lengths, 0.1999, [0.0] * 3, 0.0)
    w, rem = cp.strain_direction_and_remaining(t, cp.AxisWeightsPath((0.0, 0.0, 1.0)), lengths, 0.1, [0.0] * 3)
    assert rem == np.float32(math.log(2.0)) and [float(v) for v in w] == [0.0, 0.0, 1.0]
    t = cp.CellLengthsTarget((1.0, 2.0, 0.5))
    w, rem = cp.strain_direction_and_remaining(t, None, lengths, 0.1, [0.0] * 3)
    assert np.isclose(float(rem), math.log(2.0) + math.log(4.0), rtol=1e-6)
    assert np.allclose([float(v) for v in w], [1 / 3, 0.0, 2 / 3], rtol=1e-6)
    assert cp.target_reached(cp.CellVolumeTarget(8.0), 0.0, lengths, 0, [0] * 3, 0)
    assert cp.target_reached(cp.MeanPressureTarget(1.0), 1e-3, lengths, 0, [1.0, 2.0, 0.0], 0)
    assert not cp.target_reached(cp.DirectionalPressureTarget((0.0, 0.0, 5.0)), 1e-3, lengths, 0, [9, 9, 4], 0)
    assert cp.target_reached(cp.PenaltyEnergyTarget(2.0), 1e-3, lengths, 0, [0] * 3, 2.0)


def test_guards():
    s = tg.CompactionSettings(max_steps=3, max_pressure=10.0, max_penalty_energy=5.0)
    assert cp.guard_reached(s, 0.0, 1.0, [0.0] * 3, 0.0, 3) == "maximum compaction steps"
    assert cp.guard_reached(s, 3e-4, 1.0, [0.0] * 3, 0.0, 0) == "maximum penetration"
    assert cp.guard_reached(s, 0.0, 1.01, [0.0] * 3, 0.0, 0) == "max curvature ratio"
    assert cp.guard_reached(s, 0.0, 1.0, [0.0, 11.0, 0.0], 0.0, 0) == "maximum pressure"
    assert cp.guard_reached(s, 0.0, 1.0, [0.0] * 3, 6.0, 0) == "maximum penalty energy"
    assert cp.guard_reached(s, 0.0, 1.0, [0.0] * 3, 0.0, 0) is None


def test_settings_validation_and_text():
    s = tg.CompactionSettings()
    assert repr(s.target) == "VolumeFractionTarget(0.3)" and s.kinematics == "rigid_fiber_centers"
    assert s.weights_for(1) == cp.AxisWeightsPath((0.0, 1.0, 0.0))
    with pytest.raises(ValueError):
        tg.CompactionSettings.volume_fraction(1.0)
    with pytest.raises(ValueError):
        tg.CompactionSettings(cell_anchor=(0.5, 1.5, 0.5))
    with pytest.raises(ValueError):
        tg.CompactionSettings(min_log_strain=0.1)
    with pytest.raises(ValueError):
        tg.CompactionSettings(shrink_factor=1.0)
    with pytest.raises(ValueError):
        tg.CompactionSettings(max_pressure=0.0)
    with pytest.raises(ValueError):
        tg.CompactionSettings(kinematics="squash")
    with pytest.raises(TypeError):
        tg.CompactionSettings(target=0.3)
    with pytest.raises(ValueError):
        cp.EqualPressurePath("")
    with pytest.raises(ValueError):
        cp.DirectionalPressureTarget((0.0, 0.0, 0.0))
    assert cp.reason_text("Jammed") == "Jammed"
    assert cp.reason_text(("GuardReached", "maximum penetration")) == 'GuardReached("maximum penetration")'
    assert repr(cp.CellLengthsTarget((1e-3, 2e-3, 5e-4))) == "CellLengthsTarget(lengths=[0.001, 0.002, 0.0005])"
