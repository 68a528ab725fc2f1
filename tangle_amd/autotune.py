"""Launch-shape and grid autotuner (no counterpart in the reference, whose
launch shapes, cell size and skin are fixed).

Every choice it makes leaves the physics unchanged: the lanes per segment
in the contact pass and per fiber in the curvature sweep give bit-identical
positions, and the cell size and Verlet skin only change which extra
candidates a list holds and how often lists are rebuilt (lists hold every
pair within the skin and are summed in ascending order). It times the
candidates on a scratch world built from the same specimens, from their
generated state, and returns the fastest combination.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

from . import _native

GROUPS = ["rebuild", "contact", "measure_assess", "vertex_passes", "curvature_gs"]


@dataclass
class Tuning:
    cell_size_scale: float
    neighbor_skin_scale: float
    contact_lanes: int
    curvature_lanes: int
    trials: dict = field(default_factory=dict)

    def apply(self, world: _native.World) -> None:
        world.set_tuning(contact_lanes=self.contact_lanes, curvature_lanes=self.curvature_lanes)

    def configure(self, config: _native.Config) -> _native.Config:
        c = _native.Config.from_buffer_copy(config)
        c.cell_size_scale = self.cell_size_scale
        c.neighbor_skin_scale = self.neighbor_skin_scale
        return c


def _world(specimens, config):
    return _native.World.batch(specimens, config) if len(specimens) > 1 else _native.World(specimens[0], config)


def _group_ms(world, config, iterations):
    lib = _native.library()
    ms = (C.c_double * 5)()
    launches = (C.c_uint64 * 5)()
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    world.run_batch(config, iterations)
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    return {GROUPS[g]: ms[g] for g in range(5)}


def autotune(specimens, config: _native.Config, *, iterations: int = 32, contact_lanes=(4, 8, 16, 32),
             curvature_lanes=(4, 8, 16, 32, 64), skin_scales=(1.0, 2.0, 3.0), cell_scales=(1.0,)) -> Tuning:
    """Picks contact and curvature lanes from HIP-event group times, then
    cell size scale and skin scale from the wall time of `iterations`
    iterations (rebuilds included), each candidate starting from the
    generated state of `specimens` (PackedAssembly objects)."""
    specimens = list(specimens)
    cfg = _native.Config.from_buffer_copy(config)
    cfg.force_full_iterations = 1
    trials = {"contact": {}, "curvature": {}, "grid": {}}
    lib = _native.library()
    with _world(specimens, cfg) as world:
        initial = world.download_positions()
        _native.check(lib.tgh_world_set_timing(world._handle, 1))
        base = world.tuning()
        for lanes in contact_lanes:
            world.upload_positions(initial)
            world.set_tuning(contact_lanes=lanes, curvature_lanes=base.curvature_lanes)
            trials["contact"][lanes] = _group_ms(world, cfg, iterations)["contact"]
        best_contact = min(trials["contact"], key=trials["contact"].get)
        for lanes in curvature_lanes:
            try:
                world.set_tuning(contact_lanes=best_contact, curvature_lanes=lanes)
            except ValueError:
                continue  # the longest fiber does not fit in LDS at this width
            world.upload_positions(initial)
            trials["curvature"][lanes] = _group_ms(world, cfg, iterations)["curvature_gs"]
        best_curvature = min(trials["curvature"], key=trials["curvature"].get)
    for skin in skin_scales:
        for cell in cell_scales:
            c = _native.Config.from_buffer_copy(cfg)
            c.neighbor_skin_scale = skin
            c.cell_size_scale = cell
            with _world(specimens, c) as world:
                world.set_tuning(contact_lanes=best_contact, curvature_lanes=best_curvature)
                world.run_batch(c, iterations)  # captures the batch graph outside the timing
                world.upload_positions(initial)
                t0 = time.perf_counter()
                world.run_batch(c, iterations)
                trials["grid"][(skin, cell)] = 1000.0 * (time.perf_counter() - t0)
    skin, cell = min(trials["grid"], key=trials["grid"].get)
    return Tuning(cell_size_scale=cell, neighbor_skin_scale=skin, contact_lanes=best_contact,
                  curvature_lanes=best_curvature, trials=trials)
