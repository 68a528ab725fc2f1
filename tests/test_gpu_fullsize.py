"""The benchmark workloads at their stated sizes, through the C ABI.

- The bimodal felt: 2,000 fibers (1,763 x 7 um + 237 x 19 um, 1 mm long) in a
  1 mm periodic cube, 131,335 segments. The cell list and the neighbour pair
  sets are bit-exact against the oracle from the generated state and from a
  relaxed state, and so are the positions and every status field after 50
  full iterations.
- 256 batched specimens with fiber counts log-uniform in [50, 5,000] (10 um
  fibers of 0.5 mm, 25 segments, 10 % fill each; 289,802 fibers, 7.25 M
  segments): a sample of specimens from the whole size range follows its solo
  trajectory bit for bit inside the batch, the solo runs match the oracle
  bit for bit, and every specimen of the batch reaches its convergence gate.
"""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
import tangle_amd as tg
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def _lists_equal(gpu, ref):
    gc, go, gs = gpu.cell_list()
    oc, oo, os_ = ref.cell_list()
    assert np.array_equal(gc, oc), "cell counts differ"
    assert np.array_equal(go, oo), "cell offsets differ"
    assert np.array_equal(gs, os_), "cell segment order differs"
    gn, gl = gpu.neighbor_lists()
    on, ol = ref.neighbor_lists()
    assert np.array_equal(gn, on), "neighbour counts differ"
    ok = gn <= gl.shape[1]
    # entries past a list's count are unspecified (left from earlier builds)
    live = ok[:, None] & (np.arange(gl.shape[1])[None, :] < gn[:, None])
    assert np.array_equal(np.where(live, gl, 0), np.where(live, ol, 0)), "pair sets differ"
    return int(ok.sum()), int((~ok).sum())


def test_bimodal_full_size_lists_and_positions_bitwise(hip_device):
    print("[case] bimodal felt, 2,000 fibers: lists, 50 iterations", flush=True)
    packed = fx.bimodal_recipe()._packed(False)
    assert len(packed.counts) == 2000 and packed.vertex_count == 133_335
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    with _native.World(packed, cfg) as gpu, oracle.OracleWorld(packed, cfg) as ref:
        info = gpu.info()
        assert list(info.cells) == list(ref.info().cells) == [13, 13, 13]
        gpu.rebuild_lists()
        ref.rebuild_lists()
        listed, overflowed = _lists_equal(gpu, ref)
        print(f"[bimodal] listed {listed}, overflowed {overflowed}", flush=True)
        sg = gpu.run_batch(cfg, 50)
        so = ref.run_batch(cfg, 50)
        for field in ("total_iterations", "batch_iterations", "max_penetration", "max_displacement",
                      "max_curvature_ratio", "converged", "cell_list_overflow"):
            assert getattr(sg, field) == getattr(so, field), field
        pg, po = gpu.download_positions(), ref.download_positions()
        diff = np.nonzero(pg.view(np.uint32) != po.view(np.uint32))[0]
        assert diff.size == 0, f"{diff.size} coordinates differ, max |dx| {np.max(np.abs(pg - po)):.3e}"
        assert gpu.info().neighbor_list_rebuilds == ref.info().neighbor_list_rebuilds
        # lists rebuilt from the relaxed state, identical positions on both
        gpu.upload_positions(pg)
        ref.upload_positions(pg)
        gpu.rebuild_lists()
        ref.rebuild_lists()
        _lists_equal(gpu, ref)


def _batch256():
    recipes = fx.batched_specimen_recipes(count=256)
    return recipes, [r._packed(False) for r in recipes]


def test_batched_256_sample_follows_solo_and_oracle(hip_device):
    print("[case] 256 batched specimens: 40 iterations, sampled specimens solo and oracle", flush=True)
    recipes, packed = _batch256()
    counts = np.array([len(p.counts) for p in packed])
    assert counts.min() >= 50 and counts.max() <= 5000 and counts.sum() == 289_802
    cfg = tg.RelaxationSettings().to_config(10e-6)
    cfg.force_full_iterations = 1
    with _native.World.batch(packed, cfg) as w:
        w.run_batch(cfg, 40)
        statuses = w.specimen_status()
        together = w.specimen_positions()
    order = np.argsort(counts)
    sample = sorted({int(order[0]), int(order[-1]), int(order[len(order) // 2]), 17, 101, 202})
    for i in sample:
        print(f"[case] specimen {i}: {counts[i]} fibers", flush=True)
        with _native.World(packed[i], cfg) as solo, oracle.OracleWorld(packed[i], cfg) as ref:
            ss = solo.run_batch(cfg, 40)
            so = ref.run_batch(cfg, 40)
            alone = solo.download_positions()
            assert np.array_equal(alone.view(np.uint32), ref.download_positions().view(np.uint32)), i
            assert ss.max_penetration == so.max_penetration
        assert statuses[i].total_iterations == ss.total_iterations == 40
        assert np.array_equal(together[i].view(np.uint32), alone.view(np.uint32)), i
        assert statuses[i].max_penetration == ss.max_penetration
        assert statuses[i].max_curvature_ratio == ss.max_curvature_ratio


def test_batched_256_every_specimen_converges(hip_device):
    print("[case] 256 batched specimens to the default gate", flush=True)
    recipes, _ = _batch256()
    settings = tg.RelaxationSettings(iterations_per_batch=512)
    results = tg.relax_batch(recipes, settings, max_iterations=20000)
    iters = [r.iterations for r in results]
    print(f"[batch256] slowest {max(iters)} iterations, median {int(np.median(iters))}", flush=True)
    assert len(results) == 256
    tol = settings.to_config(10e-6).penetration_tolerance
    for i, r in enumerate(results):
        assert r.converged, (i, r.iterations, r.max_penetration)
        assert r.max_penetration <= tol
