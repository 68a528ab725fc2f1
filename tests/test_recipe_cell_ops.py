"""Recording of the cell-fit and bend-radius recipe operations
(tangle_python/src/recipe.rs:365-392, 513-543) and the host arithmetic the
cell fit runs between batches (active_capsule_bounds,
tangle_generate/src/recipe/compaction_control.rs:446-474)."""

from __future__ import annotations

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import _active_capsule_bounds, _curvature_limits


def recipe(periodic="xy"):
    cell = tg.Cell([300e-6] * 3, periodic=periodic)
    r = tg.Recipe(cell)
    pop = tg.FiberPopulation(material=tg.Material("fine", 10e-6, 100e-6), count=4, length=200e-6,
                             segments_per_fiber=4, seed=3)
    r.insert(tg.generate_fiber_population(cell, pop))
    return r


def test_fit_cell_recording():
    r = recipe()
    r.fit_cell_to_active_fibers(padding=2e-6)
    r.fit_cell_to_active_fibers(axes="z")
    assert r.operations()[-2:] == ["fit cell axes z to active fibers"] * 2
    with pytest.raises(ValueError, match="periodic cell axis 0"):
        r.fit_cell_to_active_fibers(axes="xz")
    with pytest.raises(ValueError, match="at least one"):
        r.fit_cell_to_active_fibers(axes="")
    with pytest.raises(ValueError, match="padding"):
        r.fit_cell_to_active_fibers(padding=-1.0)
    bounded = recipe(periodic="")
    bounded.fit_cell_to_active_fibers(axes="xyz")
    assert bounded.operations()[-1] == "fit cell axes xyz to active fibers"


def test_bend_radius_recording():
    r = recipe()
    r.set_min_bend_radius("fine", 1e-5)
    r.set_min_bend_radius(tg.Material("fine", 10e-6), 2.5e-4)
    assert r.operations()[-2:] == ['set "fine" min bend radius to 0.00001',
                                   'set "fine" min bend radius to 0.00025']
    for bad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(ValueError):
            r.set_min_bend_radius("fine", bad)
    with pytest.raises(ValueError):
        r.set_min_bend_radius("  ", 1e-4)
    with pytest.raises(TypeError):
        r.set_min_bend_radius(3, 1e-4)


def test_active_capsule_bounds_known_answer():
    class F:
        def __init__(self, pts, radius, step=0):
            self.placed = np.asarray(pts, dtype=np.float64)
            self.radius = radius
            self.step = step

    fibers = [F([[1.0, 1.0, 1.0], [2.0, 1.5, 1.0]], 0.1), F([[0.0, 3.0, 2.0], [0.5, 3.0, 2.5]], 0.2, step=1)]
    pos = np.concatenate([f.placed for f in fibers]).astype(np.float32)
    starts = [0, 2]
    lo, hi = _active_capsule_bounds(fibers, pos, starts, [True, False], 0.05)
    r = np.float32(np.float32(0.1) + np.float32(0.05))
    assert lo == [np.float32(1.0) - r, np.float32(1.0) - r, np.float32(1.0) - r]
    assert hi == [np.float32(2.0) + r, np.float32(1.5) + r, np.float32(1.0) + r]
    lo, hi = _active_capsule_bounds(fibers, pos, starts, [True, True], 0.0)
    assert lo[0] == np.float32(0.0) - np.float32(0.2) and hi[2] == np.float32(2.5) + np.float32(0.2)
    assert _active_capsule_bounds(fibers, pos, starts, [False, False], 0.0) == (None, None)


def test_curvature_limits():
    k = _curvature_limits([None, 1e-4, 2.0])
    assert k.dtype == np.float32
    assert k.tolist() == [0.0, float(np.float32(1e4)), 0.5]
