"""The autotuner on the GPU: its choices come from the candidate sets, a
second call on the same workload class is served from the cache, and a
world run with the tuned grid and launch shapes follows the default world
bit for bit (every choice is physics-neutral when no list overflows)."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
from tangle_amd import _native
from tangle_amd.autotune import autotune

pytestmark = pytest.mark.gpu

CONTACT, CURVATURE = (4, 8, 16), (8, 16, 32)
SKINS, CELLS = (1.0, 2.0, 3.0), (1.0, 1.5)


def _specimens():
    return [r._packed(False) for r in fx.batched_specimen_recipes(count=4, n_min=60, n_max=400)]


def test_choices_cache_and_bitwise_neutrality(hip_device, tmp_path):
    specimens = _specimens()
    cfg = fx.config(force_full_iterations=1, max_step=0.02 * 10e-6, penetration_tolerance=0.0)
    cache = tmp_path / "autotune.json"
    print("[case] autotune, cold cache", flush=True)
    t = autotune(specimens, cfg, iterations=16, contact_lanes=CONTACT, curvature_lanes=CURVATURE,
                 skin_scales=SKINS, cell_scales=CELLS, cache=cache)
    assert not t.cached
    assert t.contact_lanes in CONTACT and t.curvature_lanes in CURVATURE
    assert t.neighbor_skin_scale in SKINS and t.cell_size_scale in CELLS
    assert set(t.trials["grid"]) == {(s, c) for s in SKINS for c in CELLS}
    assert t.rebuild_interval > 0
    print("[case] autotune, warm cache", flush=True)
    again = autotune(specimens, cfg, iterations=16, contact_lanes=CONTACT, curvature_lanes=CURVATURE,
                     skin_scales=SKINS, cell_scales=CELLS, cache=cache)
    assert again.cached
    assert (again.contact_lanes, again.curvature_lanes, again.neighbor_skin_scale, again.cell_size_scale) == \
        (t.contact_lanes, t.curvature_lanes, t.neighbor_skin_scale, t.cell_size_scale)

    print("[case] tuned vs default world, 60 iterations", flush=True)
    tuned_cfg = t.configure(cfg)
    with _native.World.batch(specimens, cfg) as plain, _native.World.batch(specimens, tuned_cfg) as tuned:
        t.apply(tuned)
        plain.run_batch(cfg, 60)
        tuned.run_batch(tuned_cfg, 60)
        counts, _ = tuned.neighbor_lists()
        assert int(np.max(counts)) <= tuned.info().neighbor_block  # no list overflowed
        counts, _ = plain.neighbor_lists()
        assert int(np.max(counts)) <= plain.info().neighbor_block
        assert np.array_equal(plain.download_positions().view(np.uint32),
                              tuned.download_positions().view(np.uint32))
