"""Formation on the HIP backend against the oracle, through the C ABI:
fibers activated by formation step (kernels.rs:164-221, world.rs:2139-2180),
layer targets (kernels.rs:538-580, world.rs:1723-1790) and vertex (needle)
targets (kernels.rs:582-635, world.rs:1792-1927), with their errors
(kernels.rs:637-710).

Bars: active counts equal; cell lists and neighbour lists bit-exact over the
active segments; positions bit-exact on straight fibers (no bend limit, so no
acos) and within 1e-4 d_min on the bimodal felt; target errors equal."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def with_formation(packed, steps, layers=None):
    return _native.PackedAssembly(packed.counts, packed.placed, packed.intrinsic, packed.radius,
                                  packed.min_bend_radius, list(packed.struct.cell_origin),
                                  list(packed.struct.cell_lengths), list(packed.struct.cell_periodic),
                                  pin_fiber_ends=bool(packed.struct.pin_fiber_ends),
                                  vertex_pinned=packed.vertex_pinned, formation_steps=steps,
                                  formation_layers=layers)


def layered(packed, layers_count, axis=2):
    """Formation step and layer of every fiber from its centre height: the
    fibers of the lowest band form first."""
    pts = packed.placed.reshape(-1, 3)
    starts = np.concatenate([[0], np.cumsum(packed.counts.astype(np.int64))[:-1]])
    centers = np.array([pts[s:s + n, axis].mean() for s, n in zip(starts, packed.counts)])
    lo, hi = packed.struct.cell_origin[axis], packed.struct.cell_origin[axis] + packed.struct.cell_lengths[axis]
    band = np.minimum(((centers - lo) / (hi - lo) * layers_count).astype(np.int64), layers_count - 1)
    return band.astype(np.uint32), band.astype(np.uint32), centers


def lists_equal(gpu, ref):
    gc, go, gs = gpu.cell_list()
    oc, oo, os_ = ref.cell_list()
    assert np.array_equal(gc, oc), "cell counts differ"
    assert np.array_equal(go, oo), "cell offsets differ"
    active = int(gc.sum())
    assert np.array_equal(gs[:active], os_[:active]), "cell segment order differs"
    gn, gl = gpu.neighbor_lists()
    on, ol = ref.neighbor_lists()
    assert np.array_equal(gn, on), "neighbour counts differ"
    block = gl.shape[1]
    for s in np.nonzero(gn <= block)[0]:
        assert np.array_equal(gl[s, :gn[s]], ol[s, :gn[s]]), f"segment {s} list differs"
    return active


def test_activation_lists_bit_exact(hip_device):
    packed = fx.bimodal_recipe(scale=0.05)._packed(False)
    steps, layers, _ = layered(packed, 4)
    packed = with_formation(packed, steps, layers)
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    with _native.World(packed, cfg) as gpu, oracle.OracleWorld(packed, cfg) as ref:
        previous = 0
        for k in range(4):
            print(f"[case] activation step {k}", flush=True)
            counts = gpu.activate_formation_step(k)
            assert counts == ref.activate_formation_step(k)
            assert counts[0] > previous
            previous = counts[0]
            gpu.rebuild_lists()
            ref.rebuild_lists()
            assert lists_equal(gpu, ref) == counts[0]


def test_straight_fibers_layer_and_needle_bitwise(hip_device):
    """Straight fibers (no bend limit): activation, layer targets and needle
    targets over 3 x 30 iterations agree bit for bit."""
    base = fx.config1_recipe()._packed(False)
    steps, layers, centers = layered(base, 3)
    packed = with_formation(base, steps, layers)
    cfg = fx.hard_gate_settings(10e-6, force_full_iterations=True).to_config(10e-6)
    lower = base.struct.cell_origin[2]
    side = base.struct.cell_lengths[2]
    targets = np.float32([lower + side * (0.2 + 0.3 * k) for k in range(3)])
    with _native.World(packed, cfg) as gpu, oracle.OracleWorld(packed, cfg) as ref:
        for w in (gpu, ref):
            w.activate_formation_step(1)
            w.set_layer_targets(2, targets, first_layer=0, last_layer=1, stiffness=0.25, max_translation=2e-6)
        assert np.array_equal(gpu.download_positions().view(np.uint32), ref.download_positions().view(np.uint32))
        for _ in range(2):
            assert gpu.run_batch(cfg, 30).total_iterations == ref.run_batch(cfg, 30).total_iterations
        assert np.array_equal(gpu.download_positions().view(np.uint32), ref.download_positions().view(np.uint32))
        assert gpu.formation_target_error() == ref.formation_target_error()
        # needle the first fiber of the top layer's vertices downwards
        top = int(np.nonzero(layers == 2)[0][0])
        start = int(np.cumsum(base.counts.astype(np.int64))[top] - base.counts[top])
        needled = np.arange(start, start + int(base.counts[top]), 3, dtype=np.uint32)
        for w in (gpu, ref):
            w.activate_formation_step(2)
            w.set_vertex_targets(2, needled, -20e-6, stiffness=0.5, max_translation=1e-6)
        assert gpu.formation_target_error() == ref.formation_target_error()
        for _ in range(2):
            assert gpu.run_batch(cfg, 30).total_iterations == ref.run_batch(cfg, 30).total_iterations
        pg, po = gpu.download_positions(), ref.download_positions()
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))
        le, ve = gpu.formation_target_error()
        assert (le, ve) == ref.formation_target_error()
        assert ve is not None and ve < 20e-6
        for w in (gpu, ref):
            w.clear_vertex_targets()
            w.clear_layer_targets()
        assert gpu.formation_target_error() == (None, None)
        gpu.run_batch(cfg, 10)
        ref.run_batch(cfg, 10)
        assert np.array_equal(gpu.download_positions().view(np.uint32), ref.download_positions().view(np.uint32))


def test_bimodal_formation_positions_match(hip_device):
    """The bimodal felt (bend limits on) formed in two layers, the upper one
    placed from above while the lower one holds: positions within 1e-4 d_min."""
    base = fx.bimodal_recipe(scale=0.02)._packed(False)
    steps, layers, _ = layered(base, 2)
    packed = with_formation(base, steps, layers)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    lo, side = base.struct.cell_origin[2], base.struct.cell_lengths[2]
    with _native.World(packed, cfg) as gpu, oracle.OracleWorld(packed, cfg) as ref:
        for w in (gpu, ref):
            w.activate_formation_step(0)
            w.set_layer_targets(2, [lo + 0.25 * side], stiffness=0.25, max_translation=1e-6)
        sg, so = gpu.run_batch(cfg, 40), ref.run_batch(cfg, 40)
        assert sg.total_iterations == so.total_iterations
        for w in (gpu, ref):
            w.activate_formation_step(1)
            w.set_layer_targets(2, [lo + 0.25 * side, lo + 0.75 * side], stiffness=0.25, max_translation=1e-6)
        sg, so = gpu.run_batch(cfg, 40), ref.run_batch(cfg, 40)
        assert sg.total_iterations == so.total_iterations
        pg, po = gpu.download_positions(), ref.download_positions()
        assert float(np.max(np.abs(pg.astype(np.float64) - po))) <= 1e-4 * 7e-6
        (lg, _), (lr, _) = gpu.formation_target_error(), ref.formation_target_error()
        assert abs(lg - lr) <= 1e-4 * 7e-6


def test_formation_rejects_bad_targets(hip_device):
    packed = with_formation(fx.config1_recipe(count=4)._packed(False), [0, 0, 1, 1], [0, 0, 1, 1])
    with _native.World(packed, fx.config()) as w:
        with pytest.raises(ValueError):
            w.set_layer_targets(3, [0.0])
        with pytest.raises(ValueError):
            w.set_layer_targets(2, [0.0], first_layer=1, last_layer=0)
        with pytest.raises(ValueError):
            w.set_vertex_targets(2, [10_000], 1e-6)
        with pytest.raises(ValueError):
            w.set_vertex_targets(0, [0], 1e-6, max_translation=-1.0)
