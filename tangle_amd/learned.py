"""Learned warm start for the relaxation loop (an addition; the reference has none).

A per-segment MLP (55 -> 64 -> 64 -> 3, BF16 operands with FP32 accumulation
on the MI355X matrix cores) predicts each segment's midpoint displacement
over the next interval of the plain relaxer from its 16 nearest neighbour
capsules, its tangent and its own motion over the previous interval, in
units of that recent motion: it learns to extrapolate the relaxer's
trajectory. It is trained on the GPU, by the library's own forward and
backward kernels, on trajectories the plain relaxer generates. During a run
it proposes a displacement every `interval` iterations; a proposal that
raises the measured residual is reverted, and the plain PBD projector always
runs the later passes and judges admissibility.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native

FEATURE_WIDTH = 64


class WarmStartModel:
    """The per-segment MLP with its device-resident trajectory dataset."""

    def __init__(self, seed: int = 0):
        lib = _native.library()
        handle = C.c_void_p()
        _native.check(lib.tgh_mlp_create(int(seed), C.byref(handle)))
        self._handle = handle
        self.parameter_count = int(lib.tgh_mlp_parameter_count())

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _native.library().tgh_mlp_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters -------------------------------------------------------
    def parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count, dtype=np.float32)
        _native.check(_native.library().tgh_mlp_get_parameters(self._handle, _native.ptr(out, C.c_float), out.size))
        return out

    def set_parameters(self, values: np.ndarray) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float32)
        _native.check(_native.library().tgh_mlp_set_parameters(self._handle, _native.ptr(arr, C.c_float), arr.size))

    def save(self, path) -> None:
        np.savez(path, parameters=self.parameters())

    @classmethod
    def load(cls, path) -> "WarmStartModel":
        model = cls()
        with np.load(path, allow_pickle=False) as data:
            model.set_parameters(data["parameters"])
        return model

    # -- direct evaluation --------------------------------------------------
    def forward(self, features: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(features, dtype=np.uint16).reshape(-1, FEATURE_WIDTH)
        out = np.empty((x.shape[0], 3), dtype=np.float32)
        _native.check(_native.library().tgh_mlp_forward(self._handle, _native.ptr(x, C.c_uint16), x.shape[0],
                                                        _native.ptr(out, C.c_float)))
        return out

    def gradients(self, features: np.ndarray, targets: np.ndarray):
        x = np.ascontiguousarray(features, dtype=np.uint16).reshape(-1, FEATURE_WIDTH)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        grad = np.empty(self.parameter_count, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _native.check(_native.library().tgh_mlp_gradients(self._handle, _native.ptr(x, C.c_uint16),
                                                          _native.ptr(t, C.c_float), x.shape[0],
                                                          _native.ptr(grad, C.c_float), _native.ptr(loss, C.c_float)))
        return grad, float(loss[0])

    def time_kernels(self, rows: int, repeats: int = 10):
        """(forward ms, gradient ms) per launch over `rows` rows."""
        f, g = C.c_float(), C.c_float()
        _native.check(_native.library().tgh_mlp_time_kernels(self._handle, int(rows), int(repeats), C.byref(f),
                                                             C.byref(g)))
        return float(f.value), float(g.value)

    # -- dataset and training ---------------------------------------------
    def record_begin(self, world: _native.World) -> None:
        _native.check(_native.library().tgh_mlp_record_begin(self._handle, world._handle))

    def record_end(self, world: _native.World) -> None:
        _native.check(_native.library().tgh_mlp_record_end(self._handle, world._handle))

    @property
    def dataset_rows(self) -> int:
        return int(_native.library().tgh_mlp_dataset_rows(self._handle))

    def clear_dataset(self) -> None:
        _native.check(_native.library().tgh_mlp_dataset_clear(self._handle))

    def train(self, epochs: int = 4, batch_rows: int = 8192, learning_rate: float = 1e-3, seed: int = 0):
        losses = np.zeros(epochs, dtype=np.float32)
        _native.check(_native.library().tgh_mlp_train(self._handle, int(epochs), int(batch_rows),
                                                      float(learning_rate), int(seed),
                                                      _native.ptr(losses, C.c_float)))
        return losses


def world_features(world: _native.World) -> np.ndarray:
    """Features of every segment at the world's current positions (bf16 bits)."""
    info = world.info()
    out = np.empty((info.segment_count, FEATURE_WIDTH), dtype=np.uint16)
    _native.check(_native.library().tgh_world_features(world._handle, _native.ptr(out, C.c_uint16), out.size))
    return out


def mark_history(world: _native.World) -> None:
    """Makes the current midpoints the reference of the motion features."""
    _native.check(_native.library().tgh_world_mark_history(world._handle))


def propose(world: _native.World, model: WarmStartModel, scale: float) -> None:
    _native.check(_native.library().tgh_world_propose(world._handle, model._handle, float(scale)))


def revert(world: _native.World) -> None:
    _native.check(_native.library().tgh_world_revert_proposal(world._handle))


def revert_specimens(world: _native.World, flags) -> None:
    """Reverts the last proposal for the specimens whose flag is set."""
    arr = np.ascontiguousarray([1 if f else 0 for f in flags], dtype=np.uint32)
    _native.check(_native.library().tgh_world_revert_proposal_specimens(world._handle, _native.ptr(arr, C.c_uint32),
                                                                       arr.size))


def measure(world: _native.World, config: _native.Config) -> _native.BatchStatus:
    """Residuals of the current positions, without moving anything: a batch
    whose tolerances every state meets stops at its first (measuring) pass."""
    probe = _native.Config.from_buffer_copy(config)
    probe.penetration_tolerance = 3.0e38
    probe.curvature_ratio_tolerance = 3.0e38
    probe.force_full_iterations = 0
    return world.run_batch(probe, 1)


def collect_trajectories(model: WarmStartModel, packed: _native.PackedAssembly, config: _native.Config,
                         iterations: int, interval: int) -> int:
    """Relaxes `packed` with the plain loop for `iterations` iterations and,
    at every multiple of `interval` past the first, records one row per
    segment: the features (with the motion over the previous interval) and
    the midpoint displacement over the next interval. Returns rows added."""
    if interval <= 0 or iterations < 2 * interval:
        raise ValueError("need interval > 0 and at least two intervals of iterations")
    before = model.dataset_rows
    cfg = _native.Config.from_buffer_copy(config)
    cfg.force_full_iterations = 1
    with _native.World(packed, cfg) as world:
        mark_history(world)
        world.run_batch(cfg, interval)
        done = interval
        while done + interval <= iterations:
            model.record_begin(world)
            mark_history(world)
            world.run_batch(cfg, interval)
            model.record_end(world)
            done += interval
    return model.dataset_rows - before


@dataclass
class WarmStart:
    """How Recipe.run uses a trained model: every `interval` iterations of the
    first `active_iterations` (after a first interval that provides the
    motion history), move each vertex by `scale` times the predicted
    displacement, run `settle` plain iterations, and revert the move when the
    normalised residual is then above `tolerance` times its value before.
    `interval` should be the one the model was trained with."""

    model: WarmStartModel
    interval: int = 64
    scale: float = 1.0
    active_iterations: int = 2048
    settle: int = 8
    tolerance: float = 1.0

    def __post_init__(self):
        if self.interval <= 0 or self.active_iterations < 0 or self.settle < 0:
            raise ValueError("interval must be positive; active_iterations and settle nonnegative")
        if not (self.scale == self.scale and abs(self.scale) < 1e6) or not self.tolerance > 0:
            raise ValueError("scale must be finite and tolerance positive")


def train_warm_start(specimens, d_min: float, *, interval: int = 32, default_iterations: int = 1024,
                     cleanup_iterations: int = 1024, cleanup_preset: str | None = "contact_cleanup",
                     cleanup_after: int = 2048, epochs: int = 12, batch_rows: int = 16384,
                     learning_rate: float = 1e-3, seed: int = 0, model: WarmStartModel | None = None):
    """Trains a warm-start model on the GPU from plain-relaxer trajectories of
    `specimens` (PackedAssembly objects): `default_iterations` from each
    generated state under the default settings, then, from the state after
    `cleanup_after` default iterations, `cleanup_iterations` under the
    `cleanup_preset` overrides (the final stage of a hard-gate plan). Returns
    (model, epoch_losses)."""
    from . import RelaxationOverrides, RelaxationSettings

    model = model or WarmStartModel(seed=seed)
    base = RelaxationSettings().to_config(d_min)
    cleanup = RelaxationOverrides.preset(cleanup_preset).apply(base) if cleanup_preset else base
    for packed in specimens:
        collect_trajectories(model, packed, base, iterations=default_iterations, interval=interval)
        if cleanup_iterations <= 0:
            continue
        cfg = _native.Config.from_buffer_copy(base)
        cfg.force_full_iterations = 1
        with _native.World(packed, cfg) as world:
            world.run_batch(cfg, cleanup_after)
            warm = world.download_positions().astype(np.float64)
        s = packed.struct
        relaxed = _native.PackedAssembly(packed.counts, warm, packed.intrinsic, packed.radius, packed.min_bend_radius,
                                         [s.cell_origin[k] for k in range(3)], [s.cell_lengths[k] for k in range(3)],
                                         [bool(s.cell_periodic[k]) for k in range(3)],
                                         pin_fiber_ends=bool(s.pin_fiber_ends), vertex_pinned=packed.vertex_pinned)
        collect_trajectories(model, relaxed, cleanup, iterations=cleanup_iterations, interval=interval)
    losses = model.train(epochs=epochs, batch_rows=batch_rows, learning_rate=learning_rate, seed=seed)
    return model, losses
