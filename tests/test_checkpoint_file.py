"""Checkpoint files (CPU): round trip of the device state and recipe
progress, the atomic commit, and the rejections the reference's loader makes
(foreign magic or schema, another case id; tangle_checkpoint/src/lib.rs:114-175,
374-404)."""

from __future__ import annotations

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import checkpoint as ck
from tangle_amd._native import DeviceCheckpoint


def sample(with_targets=True):
    rng = np.random.default_rng(3)
    d = DeviceCheckpoint(1234, rng.standard_normal(30).astype(np.float32), np.array([1, 0, 1, 1], np.uint32),
                         np.zeros(30, np.float32))
    if with_targets:
        d.layer_target = dict(axis=2, first=1, last=1, stiffness=0.25, max_translation=1e-5,
                              targets=np.float32([0.1, 0.2, 0.3]))
        d.vertex_target = dict(axis=2, stiffness=1.0, max_translation=0.01, indices=np.uint32([3, 7]),
                               targets=np.float32([0.5, -0.25]), commands=np.float32([0.6, -0.2]))
    return d


@pytest.mark.parametrize("with_targets", [True, False])
def test_round_trip(tmp_path, with_targets):
    d = sample(with_targets)
    progress = {"operation": 3, "operation_begin": 1200, "events": ["a", "b"], "layer_targets": [0.5, 1.5]}
    path = tmp_path / "run.ckpt"
    size = ck.save(path, "case-1", d, progress)
    assert size == path.stat().st_size
    assert not (tmp_path / ".run.ckpt.tmp").exists()
    got, prog = ck.load(path, "case-1")
    assert prog == progress
    assert got.total_iterations == 1234
    for name in ("positions", "fiber_active", "wall_reactions"):
        assert np.array_equal(getattr(got, name).view(np.uint32), getattr(d, name).view(np.uint32))
    if with_targets:
        assert got.layer_target["first"] == 1 and np.array_equal(got.layer_target["targets"], d.layer_target["targets"])
        for key in ("indices", "targets", "commands"):
            assert np.array_equal(got.vertex_target[key], d.vertex_target[key])
        assert got.vertex_target["max_translation"] == pytest.approx(0.01)
    else:
        assert got.layer_target is None and got.vertex_target is None


def test_rejections(tmp_path):
    path = tmp_path / "run.ckpt"
    ck.save(path, "case-1", sample(), {"operation": 0})
    with pytest.raises(ck.CheckpointError, match="does not match"):
        ck.load(path, "case-2")
    data = bytearray(path.read_bytes())
    data[:8] = b"TANGLE01"
    (tmp_path / "foreign.ckpt").write_bytes(bytes(data))
    with pytest.raises(ck.CheckpointError, match="incompatible"):
        ck.load(tmp_path / "foreign.ckpt", "case-1")
    (tmp_path / "short.ckpt").write_bytes(path.read_bytes()[:-40])
    with pytest.raises(ck.CheckpointError, match="truncated"):
        ck.load(tmp_path / "short.ckpt", "case-1")
    with pytest.raises(ck.CheckpointError, match="I/O"):
        ck.load(tmp_path / "missing.ckpt", "case-1")


def test_overwrite_is_atomic(tmp_path):
    path = tmp_path / "run.ckpt"
    ck.save(path, "c", sample(), {"operation": 1})
    d = sample()
    d.total_iterations = 99
    ck.save(path, "c", d, {"operation": 2})
    got, prog = ck.load(path, "c")
    assert got.total_iterations == 99 and prog["operation"] == 2


def test_settings_validation(tmp_path):
    s = tg.CheckpointSettings("felt", tmp_path / "a.ckpt", interval_iterations=250)
    assert s.source == (tmp_path / "a.ckpt", "felt")
    s = tg.CheckpointSettings("felt", tmp_path / "a.ckpt", resume=True, resume_path=tmp_path / "b.ckpt",
                              resume_case_id="old")
    assert s.source == (tmp_path / "b.ckpt", "old")
    with pytest.raises(ValueError):
        tg.CheckpointSettings(" ", tmp_path / "a.ckpt")
    with pytest.raises(ValueError):
        tg.CheckpointSettings("felt", tmp_path / "a.ckpt", interval_iterations=0)
    with pytest.raises(ValueError):
        tg.CheckpointSettings("felt", tmp_path / "a.ckpt", resume_case_id="")


def test_named_device_states_ride_along(tmp_path):
    """Further device states (a compaction's accepted state) travel in the
    same file under their names, and a file without them loads as before."""
    path = tmp_path / "x.ckpt"
    main, extra = sample(), sample()
    extra.positions = extra.positions + np.float32(1.0)
    extra.cell_lower, extra.cell_upper = (0.0, 0.0, 0.1), (1.0, 1.0, 0.9)
    ck.save(path, "c", main, {"operation": 3}, extras={"accepted": extra})
    got, progress, extras = ck.load_with_extras(path, "c")
    assert progress == {"operation": 3}
    assert np.array_equal(got.positions, main.positions)
    assert set(extras) == {"accepted"}
    acc = extras["accepted"]
    assert np.array_equal(acc.positions, extra.positions)
    assert acc.cell_lower == (0.0, 0.0, 0.1) and acc.cell_upper == (1.0, 1.0, 0.9)
    assert acc.total_iterations == extra.total_iterations
    ck.save(path, "c", main, {"operation": 4})
    assert ck.load_with_extras(path, "c")[2] == {}
    assert ck.load(path, "c")[1] == {"operation": 4}
