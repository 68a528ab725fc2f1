"""Recipe.run budgets and checkpoints with the learned warm start on, and a
resume that restarts the formation on a compacted cell.

- A proposal's settle iterations count against the stage and the run:
  relax_for(N) ends at exactly N iterations, and no run passes its
  max_iterations.
- A checkpoint never holds a proposal the gate later rejected: a save that
  falls due inside a proposal's settle is taken after the keep/revert
  decision, so a resumed run ends as the uninterrupted one does, bit for
  bit.
- A run resumed with fresh_formation_on_resume keeps the checkpoint's
  (compacted) cell as the recipe's cell.
"""

from __future__ import annotations

import re
import shutil

import numpy as np
import pytest

import fixtures as fx
import tangle_amd as tg
from tangle_amd import checkpoint as ck
from tangle_amd import learned

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def untrained_model(hip_device):
    # an untrained network (proposals that run ahead of the relaxer)
    model = learned.WarmStartModel(seed=3)
    yield model
    model.close()


def _warm(model, **changes):
    args = dict(interval=32, scale=8.0, settle=8, active_iterations=10 ** 9)
    args.update(changes)
    return learned.WarmStart(model, **args)


def test_relax_for_with_proposals_runs_exactly_its_iterations(untrained_model):
    for n in (300, 333):
        recipe = fx.bimodal_recipe(scale=0.02)
        recipe.relax_for(n)
        res = recipe.run(tg.RelaxationSettings(max_iterations=n, iterations_per_batch=64),
                         warm_start=_warm(untrained_model))
        print(f"[budget] relax_for({n}): {res.iterations} iterations", flush=True)
        assert res.iterations == n
    # relax_until_converged within a budget too tight for the gate
    recipe = fx.bimodal_recipe(scale=0.02)
    recipe.relax_until_converged(max_iterations=5000)
    settings = tg.RelaxationSettings(penetration_tolerance=1e-12, max_iterations=250, iterations_per_batch=64)
    with pytest.raises(tg.RecipeError) as err:
        recipe.run(settings, warm_start=_warm(untrained_model))
    assert err.value.iteration == 250, str(err.value)


def _proposal_counts(events):
    for e in events:
        m = re.match(r"warm start: (\d+) proposals kept, (\d+) reverted", e)
        if m:
            return int(m.group(1)), int(m.group(2))
    return 0, 0


def test_resume_with_proposals_matches_the_uninterrupted_run(untrained_model, tmp_path, monkeypatch):
    def recipe():
        r = fx.bimodal_recipe(scale=0.02)
        r.relax_for(600)
        return r

    # lists that never overflow (an overflowed segment's sum depends on when
    # its list was built, which a restore changes)
    settings = tg.RelaxationSettings(max_iterations=600, iterations_per_batch=64, neighbor_capacity=256)
    mid = tmp_path / "mid.ckpt"
    original_save = ck.save
    saved_at = []

    def save_and_keep(path, case_id, device, progress, extras=None):
        size = original_save(path, case_id, device, progress, extras=extras)
        saved_at.append(int(device.total_iterations))
        if device.total_iterations >= 200 and not mid.exists():
            shutil.copyfile(path, mid)
        return size

    # proposals 100 times the predicted motion raise the residual and are
    # reverted: the state a checkpoint must not capture
    warm = _warm(untrained_model, scale=100.0)
    monkeypatch.setattr(ck, "save", save_and_keep)
    full = recipe().run(settings, warm_start=warm,
                        checkpoint=tg.CheckpointSettings("warm", tmp_path / "a.ckpt", interval_iterations=50))
    monkeypatch.setattr(ck, "save", original_save)
    kept, reverted = _proposal_counts(full.events)
    print(f"[resume] saves at {saved_at}; proposals kept {kept}, reverted {reverted}", flush=True)
    assert reverted > 0 and mid.exists()
    resumed = recipe().run(settings, warm_start=warm,
                           checkpoint=tg.CheckpointSettings("warm", tmp_path / "b.ckpt", interval_iterations=50,
                                                            resume=True, resume_path=mid))
    assert resumed.resumed and 200 <= resumed.resumed_iteration < full.iterations
    assert resumed.iterations == full.iterations == 600
    assert np.array_equal(bits(resumed.positions), bits(full.positions))


def test_fresh_formation_resume_keeps_the_compacted_cell(hip_device, tmp_path, monkeypatch):
    from test_gpu_compaction import bounded_bimodal_recipe

    settings = tg.RelaxationSettings(penetration_tolerance=0.05 * 7e-6, curvature_ratio_tolerance=0.01,
                                     max_iterations=400, neighbor_capacity=256)
    first = bounded_bimodal_recipe()
    first.relax_for(128)
    first.fit_cell_to_active_fibers(padding=1e-6)
    first.relax_for(200)
    mid = tmp_path / "mid.ckpt"
    original_save = ck.save

    def save_and_keep(path, case_id, device, progress, extras=None):
        size = original_save(path, case_id, device, progress, extras=extras)
        if device.total_iterations >= 200 and not mid.exists():
            shutil.copyfile(path, mid)
        return size

    monkeypatch.setattr(ck, "save", save_and_keep)
    full = first.run(settings, checkpoint=tg.CheckpointSettings("fit", tmp_path / "a.ckpt",
                                                                interval_iterations=100))
    monkeypatch.setattr(ck, "save", original_save)
    assert mid.exists()
    original = bounded_bimodal_recipe()._cell
    assert full._cell.lengths != original.lengths, "the fit did not compact the cell"
    # the same fibers, a new formation: one relaxation on the restored state
    second = bounded_bimodal_recipe()
    second.relax_for(64)
    resumed = second.run(settings, checkpoint=tg.CheckpointSettings("fit", tmp_path / "b.ckpt",
                                                                    interval_iterations=100, resume=True,
                                                                    resume_path=mid,
                                                                    fresh_formation_on_resume=True))
    assert resumed.resumed
    assert resumed._cell.origin == full._cell.origin and resumed._cell.lengths == full._cell.lengths
