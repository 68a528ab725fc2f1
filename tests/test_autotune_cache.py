"""The autotuner's persistent cache (CPU): the key follows the workload
class, the loop-shaping settings, the candidate sets, the device and the
library build; entries survive a round trip through the file; a damaged or
foreign file is ignored rather than trusted."""

from __future__ import annotations

import json

import fixtures as fx
from tangle_amd import _native
from tangle_amd import autotune as at

DEVICE = {"name": "AMD Instinct MI355X", "arch": "gfx950:sramecc+:xnack-", "compute_units": 256}
CANDIDATES = {"iterations": [32], "contact_lanes": [4, 8], "curvature_lanes": [8, 32], "skin_scales": [1.0, 2.0],
              "cell_scales": [1.0, 1.5]}


def _specimens(seed=20_260_930, count=3):
    return [r._packed(False) for r in fx.batched_specimen_recipes(count=count, n_min=50, n_max=120, seed=seed)]


def _config(**changes):
    c = fx.config()
    for k, v in changes.items():
        setattr(c, k, v)
    return c


def test_key_is_deterministic_and_follows_the_workload_class():
    a = at.workload_key(_specimens(), _config(), CANDIDATES, DEVICE, "lib1")
    assert a == at.workload_key(_specimens(), _config(), CANDIDATES, DEVICE, "lib1")
    # the same fiber counts, radii, cell and segment length from other seeds:
    # the same class (the placement seeds do not enter the key)
    sp = _specimens()
    other = [_native.PackedAssembly(p.counts, p.placed[::-1].copy(), p.intrinsic, p.radius, p.min_bend_radius,
                                    [p.struct.cell_origin[k] for k in range(3)],
                                    [p.struct.cell_lengths[k] for k in range(3)],
                                    [bool(p.struct.cell_periodic[k]) for k in range(3)]) for p in sp]
    assert at.workload_key(other, _config(), CANDIDATES, DEVICE, "lib1") == a


def test_key_changes_with_what_shapes_the_trials():
    base = at.workload_key(_specimens(), _config(), CANDIDATES, DEVICE, "lib1")
    changed = [
        at.workload_key(_specimens(count=2), _config(), CANDIDATES, DEVICE, "lib1"),
        at.workload_key(_specimens(), _config(neighbor_capacity=24), CANDIDATES, DEVICE, "lib1"),
        at.workload_key(_specimens(), _config(curvature_cleanup_sweeps=2), CANDIDATES, DEVICE, "lib1"),
        at.workload_key(_specimens(), _config(), dict(CANDIDATES, cell_scales=[1.0]), DEVICE, "lib1"),
        at.workload_key(_specimens(), _config(), CANDIDATES, dict(DEVICE, compute_units=304), "lib1"),
        at.workload_key(_specimens(), _config(), CANDIDATES, DEVICE, "lib2"),
    ]
    assert all(k != base for k in changed)
    assert len(set(changed)) == len(changed)


def test_entries_round_trip_and_coexist(tmp_path):
    path = tmp_path / "sub" / "autotune.json"
    t1 = at.Tuning(cell_size_scale=1.25, neighbor_skin_scale=2.0, contact_lanes=8, curvature_lanes=32,
                   rebuild_interval=10.7, trials={"grid": {(2.0, 1.25): 3.5}})
    t2 = at.Tuning(cell_size_scale=1.0, neighbor_skin_scale=3.0, contact_lanes=4, curvature_lanes=8)
    at.store_cached(path, "k1", t1)
    at.store_cached(path, "k2", t2)
    got1, got2 = at.load_cached(path, "k1"), at.load_cached(path, "k2")
    assert got1.cached and got2.cached
    assert (got1.cell_size_scale, got1.neighbor_skin_scale, got1.contact_lanes, got1.curvature_lanes,
            got1.rebuild_interval) == (1.25, 2.0, 8, 32, 10.7)
    assert got1.trials["grid"] == {"(2.0, 1.25)": 3.5}
    assert (got2.contact_lanes, got2.curvature_lanes) == (4, 8)
    assert at.load_cached(path, "missing") is None
    assert sorted(p.name for p in path.parent.iterdir()) == ["autotune.json"]  # no temporary left behind


def test_damaged_or_foreign_files_are_ignored(tmp_path):
    path = tmp_path / "autotune.json"
    path.write_text("{not json")
    assert at.load_cached(path, "k") is None
    path.write_text(json.dumps({"version": 999, "entries": {"k": {}}}))
    assert at.load_cached(path, "k") is None
    path.write_text(json.dumps({"version": at.CACHE_VERSION, "entries": {"k": {"contact_lanes": 4}}}))
    assert at.load_cached(path, "k") is None
    # storing over a damaged file starts a fresh one
    at.store_cached(path, "k", at.Tuning(1.0, 2.0, 4, 8))
    assert at.load_cached(path, "k").curvature_lanes == 8


def test_configure_sets_the_grid_scales_only():
    c = _config()
    t = at.Tuning(cell_size_scale=1.5, neighbor_skin_scale=3.0, contact_lanes=4, curvature_lanes=8)
    d = t.configure(c)
    assert (d.cell_size_scale, d.neighbor_skin_scale) == (1.5, 3.0)
    assert (c.cell_size_scale, c.neighbor_skin_scale) != (1.5, 3.0)  # the input is left alone
    assert d.neighbor_capacity == c.neighbor_capacity and d.max_step == c.max_step


def test_default_cache_path_honours_the_environment(tmp_path, monkeypatch):
    monkeypatch.setenv("TANGLE_AMD_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert at.default_cache_path() == tmp_path / "x.json"
    monkeypatch.delenv("TANGLE_AMD_AUTOTUNE_CACHE")
    assert at.default_cache_path().name == "autotune.json"
