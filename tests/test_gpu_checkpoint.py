"""Checkpoint and restore of the device world (DeviceWorldCheckpoint,
crates/tangle_relax/src/device/world.rs:281-304, 692-823) and of a recipe
run (crates/tangle_checkpoint/src/lib.rs), through the C ABI.

A restored world rebuilds its broad phase at the first batch, as the
reference's does. With every neighbour list within its capacity (sorted
rows, no overflow scan) the continuation is then the uninterrupted run's,
bit for bit, and, the whole loop being bit-exact, the oracle's. A segment
whose list overflowed scans its 27 cells in cell-list order; the restored
world's list is built at another moment, so its sum of corrections can round
differently (the same holds for the reference's restore)."""

from __future__ import annotations

import shutil

import numpy as np
import pytest

import fixtures as fx
import oracle
import tangle_amd as tg
from tangle_amd import _native
from tangle_amd import checkpoint as ck

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_restore_continues_bit_for_bit(hip_device):
    packed = fx.bimodal_recipe(scale=0.02)._packed(False)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True, neighbor_capacity=200).to_config(7e-6)
    with _native.World(packed, cfg) as a:
        a.run_batch(cfg, 40)
        saved = a.checkpoint()
        assert saved.total_iterations == 40
        sa = a.run_batch(cfg, 40)
        final = a.download_positions()
        counts, _ = a.neighbor_lists()
        assert counts.max() <= 200
    with _native.World.restore(packed, cfg, saved) as b:
        sb = b.run_batch(cfg, 40)
        assert sb.total_iterations == sa.total_iterations == 80
        assert np.array_equal(bits(b.download_positions()), bits(final))
    with oracle.OracleWorld(packed, cfg) as ref:
        ref.run_batch(cfg, 80)
        assert np.array_equal(bits(ref.download_positions()), bits(final))


def test_restore_keeps_activity_and_held_targets(hip_device, tmp_path):
    base = fx.config1_recipe()._packed(False)
    n = base.counts.size
    steps = np.arange(n, dtype=np.uint32) % 3
    packed = _native.PackedAssembly(base.counts, base.placed, base.intrinsic, base.radius, None,
                                    list(base.struct.cell_origin), list(base.struct.cell_lengths),
                                    list(base.struct.cell_periodic), formation_steps=steps, formation_layers=steps)
    cfg = fx.hard_gate_settings(10e-6, force_full_iterations=True).to_config(10e-6)
    lo, side = base.struct.cell_origin[2], base.struct.cell_lengths[2]
    with _native.World(packed, cfg) as a:
        a.activate_formation_step(1)
        a.set_layer_targets(2, [lo + 0.3 * side, lo + 0.6 * side], stiffness=0.25, max_translation=2e-6)
        a.set_vertex_targets(2, [5, 40, 77], -15e-6, stiffness=0.5, max_translation=1e-6)
        a.run_batch(cfg, 20)
        saved = a.checkpoint()
        assert saved.layer_target is not None and saved.vertex_target is not None
        assert saved.fiber_active.tolist() == [1 if s <= 1 else 0 for s in steps]
        a.run_batch(cfg, 20)
        final, errors = a.download_positions(), a.formation_target_error()
    # through a file, as a resumed run reads it
    path = tmp_path / "restore_test.ckpt"
    ck.save(path, "restore-test", saved, {"operation": 0})
    loaded, _ = ck.load(path, "restore-test")
    with _native.World.restore(packed, cfg, loaded) as b:
        assert b.formation_target_error() != (None, None)
        b.run_batch(cfg, 20)
        assert np.array_equal(bits(b.download_positions()), bits(final))
        assert b.formation_target_error() == errors
        # fibers of step 2 stayed out, untouched
        p = b.download_positions().reshape(-1, 3)
        starts = np.concatenate([[0], np.cumsum(base.counts)[:-1]])
        for f in np.nonzero(steps == 2)[0][:5]:
            s0, c = int(starts[f]), int(base.counts[f])
            assert np.array_equal(bits(p[s0:s0 + c]), bits(base.placed.reshape(-1, 3)[s0:s0 + c]))


def test_restore_rejects_another_assembly(hip_device):
    packed = fx.bimodal_recipe(scale=0.02)._packed(False)
    other = fx.config1_recipe()._packed(False)
    cfg = fx.config()
    with _native.World(packed, cfg) as a:
        saved = a.checkpoint()
    with pytest.raises(ValueError):
        _native.World.restore(other, cfg, saved)


def test_recipe_resume_matches_the_uninterrupted_run(hip_device, tmp_path, monkeypatch):
    from test_gpu_formation import needled_felt_toy
    recipe, settings = needled_felt_toy(layer_count=3)
    # lists that never overflow: an overflowed segment scans its 27 cells in
    # the order of the cell list it was built with, which a restore rebuilds
    # at another moment (so sums of its corrections may round differently)
    settings = settings.replace(neighbor_capacity=256)
    mid = tmp_path / "mid.ckpt"
    original_save = ck.save

    def save_and_keep(path, case_id, device, progress, extras=None):
        size = original_save(path, case_id, device, progress, extras=extras)
        if device.total_iterations >= 400 and not mid.exists():
            shutil.copyfile(path, mid)
        return size

    monkeypatch.setattr(ck, "save", save_and_keep)
    settings_a = tg.CheckpointSettings("felt-toy", tmp_path / "run.ckpt", interval_iterations=100)
    full = recipe.run(settings, checkpoint=settings_a)
    monkeypatch.setattr(ck, "save", original_save)
    assert full.checkpoint_saves >= 5 and full.last_checkpoint_iteration == full.iterations
    assert mid.exists()
    recipe_b, _ = needled_felt_toy(layer_count=3)
    settings_b = tg.CheckpointSettings("felt-toy", tmp_path / "run_b.ckpt", interval_iterations=100, resume=True,
                                       resume_path=mid)
    resumed = recipe_b.run(settings, checkpoint=settings_b)
    print(f"[resume] resumed at {resumed.resumed_iteration}, iterations {resumed.iterations} vs {full.iterations}",
          flush=True)
    assert resumed.resumed and 400 <= resumed.resumed_iteration < full.iterations
    assert resumed.iterations == full.iterations
    assert resumed.needling == full.needling
    assert resumed.events == full.events
    assert np.array_equal(bits(resumed.positions), bits(full.positions))
    # another recipe's checkpoint is refused
    other = tg.Recipe(tg.Cell([1.0, 1.0, 0.32]))
    other.insert(tg.FiberCollection.from_centerlines([[[0.1, 0.5, 0.1], [0.9, 0.5, 0.1]]], tg.Material("f", 0.037)))
    with pytest.raises(tg.CheckpointError):
        other.run(settings, checkpoint=settings_b)


def test_resume_keeps_cell_bend_radii_and_junctions(hip_device, tmp_path, monkeypatch):
    """A recipe that fits the cell, changes a bend radius and captures
    junctions, resumed from a checkpoint taken after those operations, ends
    as the uninterrupted run does: same iterations, events, junctions, cell
    and positions bit for bit."""
    from test_gpu_compaction import bounded_bimodal_recipe

    def recipe():
        r = bounded_bimodal_recipe()
        r.relax_for(128)
        r.set_min_bend_radius("coarse", 120e-6)
        r.fit_cell_to_active_fibers(padding=1e-6)
        r.capture_junctions(tg.JunctionPolicy("bonds", max_surface_gap=0.5e-6))
        r.relax_for(256)
        r.capture_junctions(tg.JunctionPolicy("late", max_surface_gap=0.2e-6))
        return r

    settings = tg.RelaxationSettings(penetration_tolerance=0.05 * 7e-6, curvature_ratio_tolerance=0.01,
                                     max_iterations=700, neighbor_capacity=256)
    mid = tmp_path / "mid.ckpt"
    original_save = ck.save

    def save_and_keep(path, case_id, device, progress, extras=None):
        size = original_save(path, case_id, device, progress, extras=extras)
        if device.total_iterations >= 200 and not mid.exists():
            shutil.copyfile(path, mid)
        return size

    monkeypatch.setattr(ck, "save", save_and_keep)
    full = recipe().run(settings, checkpoint=tg.CheckpointSettings("cells", tmp_path / "a.ckpt",
                                                                   interval_iterations=100))
    monkeypatch.setattr(ck, "save", original_save)
    assert mid.exists() and full.junction_count > 0
    resumed = recipe().run(settings, checkpoint=tg.CheckpointSettings("cells", tmp_path / "b.ckpt",
                                                                      interval_iterations=100, resume=True,
                                                                      resume_path=mid))
    print(f"[resume] at {resumed.resumed_iteration}: {resumed.iterations} vs {full.iterations} iterations",
          flush=True)
    assert resumed.resumed and 200 <= resumed.resumed_iteration < full.iterations
    assert resumed.iterations == full.iterations
    assert resumed.events == full.events
    assert resumed.junctions == full.junctions
    assert resumed._cell.origin == full._cell.origin and resumed._cell.lengths == full._cell.lengths
    assert np.array_equal(bits(resumed.positions), bits(full.positions))


def test_resume_inside_a_compaction_matches_the_uninterrupted_run(hip_device, tmp_path, monkeypatch):
    """A checkpoint taken while a closed-loop compaction is in progress holds
    the controller's state (strain schedule, windows, steps, work, last
    metrics, the accepted state it would roll back to): the resumed run ends
    as the uninterrupted one does, with the same compaction and step
    reports, events, cell and positions bit for bit."""
    from test_gpu_compaction import bounded_bimodal_recipe
    from tangle_amd import compaction as cp

    def recipe():
        r = bounded_bimodal_recipe()
        r.relax_for(128)
        solid = cp.nominal_fiber_volume(r._fibers, None)
        side = r._cell.lengths
        target = round(solid / (side[0] * side[1] * side[2]) * 1.12, 4)
        r.compact(tg.CompactionSettings.volume_fraction(target, kinematics="moving_walls", max_curvature_ratio=1.05))
        r.relax_for(64)
        return r

    settings = tg.RelaxationSettings(penetration_tolerance=0.05 * 7e-6, curvature_ratio_tolerance=0.01,
                                     max_iterations=6000, neighbor_capacity=256)
    saved = []
    original_save = ck.save

    def save_and_keep(path, case_id, device, progress, extras=None):
        size = original_save(path, case_id, device, progress, extras=extras)
        state = progress.get("compaction_state")
        # keep the first save taken after the compaction accepted a step
        if state is not None and state["steps"] >= 1 and not saved:
            keep = tmp_path / "mid.ckpt"
            shutil.copyfile(path, keep)
            saved.append((keep, device.total_iterations, state["steps"]))
        return size

    monkeypatch.setattr(ck, "save", save_and_keep)
    full = recipe().run(settings, checkpoint=tg.CheckpointSettings("compact", tmp_path / "a.ckpt",
                                                                   interval_iterations=50))
    monkeypatch.setattr(ck, "save", original_save)
    assert saved, "no checkpoint fell inside the compaction"
    mid, at, steps = saved[0]
    assert len(full.compactions) == 1 and full.compactions[0]["steps"] > steps
    print(f"[resume] inside the compaction at iteration {at} after {steps} steps", flush=True)
    resumed = recipe().run(settings, checkpoint=tg.CheckpointSettings("compact", tmp_path / "b.ckpt",
                                                                      interval_iterations=50, resume=True,
                                                                      resume_path=mid))
    assert resumed.resumed and resumed.resumed_iteration == at
    assert resumed.iterations == full.iterations
    assert resumed.compactions == full.compactions
    assert resumed.compaction_steps == full.compaction_steps
    assert resumed.events == full.events
    assert resumed._cell.origin == full._cell.origin and resumed._cell.lengths == full._cell.lengths
    assert np.array_equal(bits(resumed.positions), bits(full.positions))
