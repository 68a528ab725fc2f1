"""Formation targets in the oracle, pinned by closed forms (CPU): the layer
target translation law clamp((target - mean) * stiffness, +-max) applied at
install and after every iteration (kernels.rs:538-580, world.rs:1420-1482,
1752-1803), the needle command law (kernels.rs:582-635), activation by
formation step (kernels.rs:164-221) and the target errors (kernels.rs:637-710)."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native


def straight(steps, layers, z=1.0, n=5):
    fiber = np.array([[0.5 + 0.2 * i, 1.0, z] for i in range(n)])
    return _native.PackedAssembly([n], fiber, fiber, [0.05], None, [0, 0, 0], [4.0] * 3, [True, True, False],
                                  formation_steps=steps, formation_layers=layers)


def quiet_config():
    return fx.config(force_full_iterations=1, max_step=0.01, penetration_tolerance=1e-6)


def test_layer_target_translation_law():
    packed = straight(steps=[0], layers=[0])
    cfg = quiet_config()
    f32 = np.float32
    with oracle.OracleWorld(packed, cfg) as w:
        w.set_layer_targets(2, [1.5], stiffness=0.5, max_translation=0.1)
        z = w.download_positions().reshape(-1, 3)[:, 2]
        assert np.all(z == f32(1.0) + f32(0.1))  # the install applies once: (1.5 - 1.0) * 0.5 clamps to 0.1
        expect = f32(1.1)
        for _ in range(6):  # one application per iteration
            t = np.clip((f32(1.5) - expect) * f32(0.5), f32(-0.1), f32(0.1))
            expect = f32(expect + t)
        w.run_batch(cfg, 6)
        z = w.download_positions().reshape(-1, 3)[:, 2]
        assert np.all(np.abs(z - expect) <= 4e-7)
        layer, _ = w.formation_target_error()
        assert layer == pytest.approx(1.5 - float(expect), abs=1e-6)
        w.clear_layer_targets()
        w.run_batch(cfg, 3)
        assert np.all(np.abs(w.download_positions().reshape(-1, 3)[:, 2] - expect) <= 4e-7)


def test_layer_targets_skip_other_layers_and_unlayered_fibers():
    fibers = [[[0.5 + 0.2 * i, y, 1.0] for i in range(4)] for y in (0.5, 1.5, 2.5)]
    placed = np.concatenate([np.array(f) for f in fibers])
    packed = _native.PackedAssembly([4, 4, 4], placed, placed, [0.05] * 3, None, [0, 0, 0], [4.0] * 3,
                                    [True, True, False], formation_layers=[0, 1, 0xFFFFFFFF])
    with oracle.OracleWorld(packed, quiet_config()) as w:
        w.set_layer_targets(2, [1.2, 3.0], first_layer=1, last_layer=1, stiffness=1.0, max_translation=5.0)
        z = w.download_positions().reshape(3, 4, 3)[:, :, 2]
    assert np.all(z[0] == np.float32(1.0)) and np.all(z[2] == np.float32(1.0))
    assert np.all(z[1] == np.float32(3.0))


def test_needle_command_law():
    packed = straight(steps=[0], layers=[0xFFFFFFFF])
    cfg = quiet_config()
    with oracle.OracleWorld(packed, cfg) as w:
        w.set_vertex_targets(2, [2], -0.3, stiffness=1.0, max_translation=0.1)
        _, err = w.formation_target_error()
        assert err == pytest.approx(0.3, abs=1e-6)
        w.run_batch(cfg, 2)
        z = w.download_positions().reshape(-1, 3)[:, 2]
        assert z[2] == np.float32(np.float32(np.float32(1.0) - np.float32(0.1)) - np.float32(0.1))
        w.run_batch(cfg, 5)
        _, err = w.formation_target_error()
        assert err <= 1e-6  # the command reached the target and holds it


def test_activation_by_formation_step():
    # two crossing fibers: only the first is active at step 0, so nothing overlaps
    a = [[0.2 + 0.2 * i, 1.0, 1.0] for i in range(9)]
    b = [[1.0, 0.2 + 0.2 * i, 1.05] for i in range(9)]
    placed = np.concatenate([np.array(a), np.array(b)])
    packed = _native.PackedAssembly([9, 9], placed, placed, [0.1, 0.1], None, [0, 0, 0], [2.0] * 3, [True] * 3,
                                    formation_steps=[0, 1])
    cfg = fx.config(penetration_tolerance=1e-5, correction_fraction=1.0, max_step=0.2)
    with oracle.OracleWorld(packed, cfg) as w:
        assert w.activate_formation_step(0) == (8, 9)
        st = w.run_batch(cfg, 5)
        assert st.converged and st.max_penetration == 0.0
        pos = w.download_positions().reshape(-1, 3)
        assert np.array_equal(pos[9:], placed[9:].astype(np.float32))  # the inactive fiber never moved
        counts, offsets, segs = w.cell_list()
        assert int(counts.sum()) == 8 and set(segs[:8].tolist()) == set(range(8))
        assert w.activate_formation_step(1) == (16, 18)
        w.run_batch(cfg, 1)
        pos = w.download_positions().reshape(-1, 3)
        assert pos[13, 2] > np.float32(1.05)  # the crossing is found once both are active
        counts, _, _ = w.cell_list()
        assert int(counts.sum()) == 16
        st = w.run_batch(cfg, 20)
        assert st.converged
