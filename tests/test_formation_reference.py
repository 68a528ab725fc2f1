"""The reference's own formation world tests as known answers, on the oracle
(CPU) and on the HIP world (GPU): crates/tangle_relax/src/device/world_tests.rs,
cited per test. The reference runs them with `upload(..., max_step)`; the
fourth argument there is this config's `max_step`."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("hip", id="hip", marks=pytest.mark.gpu)]
NONE = 0xFFFFFFFF


@pytest.fixture
def make_world(request):
    def make(kind, packed, cfg):
        if kind == "hip":
            request.getfixturevalue("hip_device")
            return _native.World(packed, cfg)
        return oracle.OracleWorld(packed, cfg)
    return make


def two_layer(z, layers, steps=None, cube=1.0):
    """Two one-segment fibers (rest [0,0,0]-[0.2,0,0]) at heights z, r = 0.01."""
    placed = np.array([[0.4, 0.5, z[0]], [0.6, 0.5, z[0]], [0.4, 0.5, z[1]], [0.6, 0.5, z[1]]])
    rest = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]] * 2)
    return _native.PackedAssembly([2, 2], placed, rest, [0.01, 0.01], None, [0, 0, 0], [cube] * 3, [False] * 3,
                                  formation_steps=steps, formation_layers=layers)


@pytest.mark.parametrize("kind", BACKENDS)
def test_formation_steps_activate_prepacked_fibers(make_world, kind):
    # world_tests.rs:47-75
    base = fx.crossed()
    packed = _native.PackedAssembly(base.counts, base.placed, base.intrinsic, base.radius, None,
                                    [0, 0, 0], [4.0] * 3, [False] * 3, formation_steps=[0, 1])
    with make_world(kind, packed, fx.config(max_step=0.2)) as w:
        assert w.activate_formation_step(0) == (1, 2)
        assert w.activate_formation_step(1) == (2, 4)


@pytest.mark.parametrize("kind", BACKENDS)
def test_layer_target_command_updates_resident_geometry(make_world, kind):
    # world_tests.rs:556-587
    with make_world(kind, two_layer((0.2, 0.8), [0, 1]), fx.config(max_step=0.01)) as w:
        w.set_layer_targets(2, [0.3, 0.7], stiffness=1.0, max_translation=1.0)
        z = w.download_positions().reshape(-1, 3)[:, 2]
    assert np.all(np.abs(z[:2] - 0.3) < 1e-6) and np.all(np.abs(z[2:] - 0.7) < 1e-6)


@pytest.mark.parametrize("kind", BACKENDS)
def test_single_layer_target_leaves_the_relaxed_stack_free(make_world, kind):
    # world_tests.rs:589-624 (apply_single_layer_target: first = last = layer)
    with make_world(kind, two_layer((0.24, 0.8), [0, 1]), fx.config(max_step=0.01)) as w:
        w.set_layer_targets(2, [0.2, 0.7], first_layer=1, last_layer=1, stiffness=1.0, max_translation=1.0)
        layer, vertex = w.formation_target_error()
        assert vertex is None and layer < 1e-6
        z = w.download_positions().reshape(-1, 3)[:, 2]
    assert np.all(np.abs(z[:2] - 0.24) < 1e-6) and np.all(np.abs(z[2:] - 0.7) < 1e-6)


@pytest.mark.parametrize("kind", BACKENDS)
def test_layer_targets_do_not_move_inactive_future_layers(make_world, kind):
    # world_tests.rs:626-658
    with make_world(kind, two_layer((0.2, 0.8), [0, 1], steps=[0, 1]), fx.config(max_step=0.01)) as w:
        w.activate_formation_step(0)
        w.set_layer_targets(2, [0.3, 0.7], stiffness=1.0, max_translation=1.0)
        z = w.download_positions().reshape(-1, 3)[:, 2]
    assert abs(z[0] - 0.3) < 1e-6 and abs(z[2] - 0.8) < 1e-6


@pytest.mark.parametrize("kind", BACKENDS)
def test_vertex_target_is_initialized_from_resident_geometry(make_world, kind):
    # world_tests.rs:660-696. The reference runs it under rigid-translation
    # motion; under the flexible model the one iteration also leaves the
    # untargeted vertices in place (the target is written after the
    # constraint passes, which found the straight fiber at rest).
    placed = np.array([[0.4, 0.5, 0.5], [0.5, 0.5, 0.5], [0.6, 0.5, 0.5]])
    rest = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.2, 0.0, 0.0]])
    packed = _native.PackedAssembly([3], placed, rest, [0.01], None, [0, 0, 0], [1.0] * 3, [False] * 3)
    cfg = fx.config(max_step=0.1, force_full_iterations=1)
    with make_world(kind, packed, cfg) as w:
        w.set_vertex_targets(2, [1], -0.1, stiffness=1.0, max_translation=0.2)
        layer, vertex = w.formation_target_error()
        assert layer is None and abs(vertex - 0.1) < 1e-6
        w.run_batch(cfg, 1)
        assert abs(w.formation_target_error()[1]) < 1e-6
        p = w.download_positions()
    assert abs(p[5] - 0.4) < 1e-6 and abs(p[2] - 0.5) < 1e-6 and abs(p[8] - 0.5) < 1e-6


@pytest.mark.parametrize("kind", BACKENDS)
def test_vertex_command_advances_independently_of_mechanics(make_world, kind):
    # world_tests.rs:698-744: the command walks 0.01 per iteration toward the
    # target 0.95 whatever the constraints do; after 3 iterations the vertex
    # sits at the command 0.97 and the residual is 0.02.
    placed = np.array([[0.8, 1.0, 1.0], [1.0, 1.0, 1.0], [1.2, 1.0, 1.0]])
    rest = np.array([[-0.15, 0.0, 0.0], [0.0, 0.0, 0.0], [0.15, 0.0, 0.0]])
    packed = _native.PackedAssembly([3], placed, rest, [0.01], [0.25], [0, 0, 0], [2.0] * 3, [False] * 3,
                                    pin_fiber_ends=True)
    cfg = fx.config(max_step=0.02, force_full_iterations=1, stretch_stiffness=1.0, bend_stiffness=1.0,
                    curvature_limit_stiffness=1.0, constraint_iterations=4)
    with make_world(kind, packed, cfg) as w:
        w.set_vertex_targets(1, [1], -0.05, stiffness=1.0, max_translation=0.01)
        w.run_batch(cfg, 3)
        p = w.download_positions()
        assert abs(p[4] - 0.97) < 1e-6
        assert abs(w.formation_target_error()[1] - 0.02) < 1e-6


@pytest.mark.parametrize("kind", BACKENDS)
def test_persistent_layer_targets_are_independent_of_batch_boundaries(make_world, kind):
    # world_tests.rs:746-785
    placed = np.array([[0.4, 0.5, 0.2], [0.6, 0.5, 0.2]])
    rest = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]])
    packed = _native.PackedAssembly([2], placed, rest, [0.01], None, [0, 0, 0], [1.0] * 3, [False] * 3,
                                    formation_layers=[0])
    cfg = fx.config(max_step=0.01, force_full_iterations=1)

    def run(batches):
        with make_world(kind, packed, cfg) as w:
            w.set_layer_targets(2, [0.8], stiffness=0.2, max_translation=1.0)
            for n in batches:
                w.run_batch(cfg, n)
            return w.download_positions()

    assert np.array_equal(run([2, 3]), run([5]))
