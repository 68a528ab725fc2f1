"""Batched specimens (no counterpart in the reference): several independent
specimens, each with its own cell, grid, skin, convergence gate and
neighbour-list rebuild flag, relaxed by the same launches. A specimen in a
batch must follow exactly the trajectory it follows alone: its lists are
rebuilt when its own vertices have moved half its skin, so even an overflowed
list (summed in stencil order, which depends on when it was built) matches,
and every other kernel is per vertex or per fiber."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import tangle_amd as tg
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def _specimens():
    # periodic straight fibers, a walled dense mat, a pinned zigzag with a bend
    # limit, and a bimodal felt with bend limits: different grids, walls,
    # pins and curvature work in one world
    return [fx.config1_recipe(count=12)._packed(False), fx.dense_mat(), fx.zigzag(),
            fx.bimodal_recipe(scale=0.02)._packed(False)]


def _config(**changes):
    cfg = fx.config(penetration_tolerance=1e-9, max_step=0.01 * 10e-6, neighbor_capacity=64,
                    curvature_ratio_tolerance=1e-5)
    for k, v in changes.items():
        setattr(cfg, k, v)
    return cfg


def _alone(packed, cfg, iterations):
    with _native.World(packed, cfg) as w:
        st = w.run_batch(cfg, iterations)
        counts, _ = w.neighbor_lists()
        return w.download_positions(), st, int(counts.max()), int(w.info().neighbor_block)


def test_batched_specimens_follow_their_solo_trajectories(hip_device):
    specimens = _specimens()
    cfg = _config(force_full_iterations=1)
    with _native.World.batch(specimens, cfg) as w:
        assert w.info().specimen_count == len(specimens)
        w.run_batch(cfg, 40)
        statuses = w.specimen_status()
        together = w.specimen_positions()
    for i, packed in enumerate(specimens):
        print(f"[case] specimen {i}", flush=True)
        alone, st, max_count, block = _alone(packed, cfg, 40)
        assert statuses[i].total_iterations == st.total_iterations == 40
        assert np.array_equal(together[i].view(np.uint32), alone.view(np.uint32)), i
        assert statuses[i].max_penetration == st.max_penetration
        assert statuses[i].max_curvature_ratio == st.max_curvature_ratio


def test_overflowed_lists_follow_their_solo_trajectories(hip_device):
    """Four-entry lists overflow in the dense specimens, whose contact sums
    then run in stencil-scan order, and the specimens rebuild at different
    iterations: each must still match its solo run bit for bit."""
    specimens = _specimens()
    cfg = _config(force_full_iterations=1, neighbor_capacity=4)
    with _native.World.batch(specimens, cfg) as w:
        w.run_batch(cfg, 60)
        together = w.specimen_positions()
        batch_rebuilds = w.info().neighbor_list_rebuilds
    overflowed, solo_rebuilds = 0, []
    for i, packed in enumerate(specimens):
        with _native.World(packed, cfg) as one:
            one.run_batch(cfg, 60)
            counts, _ = one.neighbor_lists()
            alone = one.download_positions()
            info = one.info()
            solo_rebuilds.append(info.neighbor_list_rebuilds)
            overflowed += int((counts > info.neighbor_block).any())
        print(f"[case] specimen {i}: solo rebuilds {solo_rebuilds[-1]}", flush=True)
        assert np.array_equal(together[i].view(np.uint32), alone.view(np.uint32)), i
    assert overflowed >= 2, "the case must exercise overflowed lists"
    # the batch rebuilt in every pass in which some specimen did
    assert max(solo_rebuilds) <= batch_rebuilds <= sum(solo_rebuilds)
    assert len(set(solo_rebuilds)) > 1, "the specimens must rebuild at different times"


def test_each_specimen_stops_on_its_own_gate(hip_device):
    specimens = _specimens()
    cfg = _config(penetration_tolerance=1e-3 * 10e-6)
    with _native.World.batch(specimens, cfg) as w:
        for _ in range(10):
            agg = w.run_batch(cfg, 200)
            if agg.converged:
                break
        statuses = w.specimen_status()
        together = w.specimen_positions()
    for i, packed in enumerate(specimens):
        alone_world = _native.World(packed, cfg)
        with alone_world:
            for _ in range(10):
                st = alone_world.run_batch(cfg, 200)
                if st.converged:
                    break
            alone = alone_world.download_positions()
        print(f"[case] specimen {i}: {statuses[i].total_iterations} vs {st.total_iterations}", flush=True)
        assert statuses[i].total_iterations == st.total_iterations
        assert bool(statuses[i].converged) == bool(st.converged)
        assert np.array_equal(together[i].view(np.uint32), alone.view(np.uint32)), i


def test_relax_batch_facade(hip_device):
    recipes = fx.batched_specimen_recipes(count=6, n_min=20, n_max=200)
    results = tg.relax_batch(recipes, tg.RelaxationSettings(max_iterations=4000))
    assert len(results) == 6
    for r, recipe in zip(results, recipes):
        assert r.converged, r
        assert r.positions.size == 3 * sum(len(f.placed) for f in recipe._fibers)
        assert r.max_penetration <= 0.01 * 10e-6


def test_batched_warm_start_reaches_the_gate_in_fewer_iterations(hip_device):
    """The learned warm start on a batched world: each specimen still
    relaxing takes the proposals and keeps them only when its own residual
    does not rise. 32 specimens reach penetration <= 1e-3 d with every
    specimen converged, the slowest in well under half the plain
    iterations (measured on 64 specimens: 8,788 -> 1,436)."""
    import tangle_amd as tg
    from tangle_amd import learned

    d = 10e-6
    train = [r._packed(False) for r in fx.batched_specimen_recipes(count=2, n_min=1500, n_max=2000, seed=777)]
    model, losses = learned.train_warm_start(train, d, interval=32, default_iterations=1024, cleanup_iterations=0,
                                             epochs=12, batch_rows=16384)
    assert losses[-1] < losses[0]
    out = {}
    for label in ("plain", "learned"):
        recipes = fx.batched_specimen_recipes(count=32, n_max=2000)
        settings = tg.RelaxationSettings(penetration_tolerance=1e-3 * d, curvature_ratio_tolerance=1e-5)
        ws = (learned.WarmStart(model, interval=32, scale=8.0, settle=4, active_iterations=10 ** 9)
              if label == "learned" else None)
        res = tg.relax_batch(recipes, settings, max_iterations=20000, warm_start=ws)
        assert all(r.converged and r.max_penetration <= 1e-3 * d for r in res), label
        out[label] = max(r.iterations for r in res)
        print(f"[batched warm start] {label}: slowest specimen {out[label]} iterations", flush=True)
    model.close()
    assert out["learned"] <= 0.5 * out["plain"], out
