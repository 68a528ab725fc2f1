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
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))
        assert gpu.formation_target_error() == ref.formation_target_error()


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


def needled_felt_toy(layer_count=6):
    """The reference's six-layer needled felt toy
    (crates/tangle_python/python/examples/native/needled_felt_toy.py:14-84)
    up to its final compaction, which is out of this scope: plies placed a
    gap above the last one and needled through the stack."""
    import math
    import tangle_amd as tg
    gap = 0.040
    fiber = tg.Material("needled felt fiber", diameter=0.037, min_bend_radius=0.050)
    pop = tg.FiberPopulation(material=fiber, count=180, segments_per_fiber=10, seed=20_260_929, length=(0.36, 0.48),
                             diameter=(0.032, 0.042), curvature_amplitude=(0.0, 0.006),
                             orientation=tg.PlanarOrientation(max_tilt=math.radians(6.0)),
                             position=tg.LayeredPosition(layer_count, jitter_fraction=0.12),
                             max_attempts_per_fiber=1_024)
    cell = tg.Cell([1.0, 1.0, 0.32])
    population = tg.generate_fiber_population(cell, pop, name="batting")
    recipe = tg.Recipe(cell)
    recipe.insert(population.select_layer(0))
    recipe.insert(population.select_layer(1))
    recipe.relax_for(80)
    recipe.place_layer_above(1, gap=gap, stiffness=1.0, max_translation=0.08)
    recipe.relax_for(120)
    for layer in range(2, layer_count):
        recipe.insert(population.select_layer(layer))
        with recipe.place_layer_above(layer, gap=gap, stiffness=1.0, max_translation=0.08):
            recipe.relax_for(100)
        with recipe.needle_layer(layer, footprint=tg.RandomFiberFraction(0.30, seed=20_260_930 + layer),
                                 depth=2.0 * gap, stiffness=1.0, max_translation=0.012,
                                 max_translation_over_diameter=0.25):
            recipe.relax_for(160)
        recipe.relax_for(40)
    settings = tg.RelaxationSettings(constraint_iterations=4, max_iterations=24_000, iterations_per_batch=20,
                                     penetration_tolerance=5.0e-4, max_step=0.005)
    return recipe, settings


def test_needled_felt_toy_recipe(hip_device, tmp_path):
    recipe, settings = needled_felt_toy()
    result = recipe.run(settings)
    # the reference example writes its result for OVITO
    result.write_ovito(tmp_path / "final.dump", view_script_path=tmp_path / "view.py", coloring="curvature_ratio")
    dump = (tmp_path / "final.dump").read_text().splitlines()
    assert dump[3] == str(int(result.vertex_counts.sum()) - result.fiber_count)
    first = dump[9].split()
    p = result.positions.reshape(-1, 3).astype(np.float64)
    assert np.allclose([float(v) for v in first[10:13]], 0.5 * (p[0] + p[1]), rtol=0, atol=1e-15)
    print(f"[felt toy] iterations={result.iterations} post={result.post_recipe_iterations} "
          f"pen={result.max_penetration:.3e} curv={result.max_curvature_ratio:.5f} needling={result.needling} "
          f"wall={result.wall_seconds:.2f}s", flush=True)
    fixed = 80 + 120 + 4 * (100 + 160 + 40)
    assert result.iterations - result.post_recipe_iterations == fixed
    assert len(result.needling) == 4 and all(n[1] > 0 for n in result.needling)
    assert result.converged and result.max_penetration <= 5.0e-4
    assert sum(e.startswith("needle ") for e in result.events) == 4
    # the plies end stacked: each layer's mean height above the last
    z = [np.mean([p[2] for p in line]) for line in result.centerlines()]
    layers = [f.formation_layer for f in recipe._fibers]
    means = [np.mean([zz for zz, l in zip(z, layers) if l == k]) for k in range(6)]
    assert all(b > a for a, b in zip(means, means[1:]))


def test_needles_pull_their_vertices_by_depth(hip_device):
    """One needle command held until settled moves every selected vertex by
    the commanded depth against the stack axis (recipe.rs:1091-1177)."""
    import tangle_amd as tg
    settings = tg.RelaxationSettings(constraint_iterations=4, max_iterations=4_000, iterations_per_batch=20,
                                     penetration_tolerance=5.0e-4, max_step=0.005)
    recipe = tg.Recipe(tg.Cell([1.0, 1.0, 0.32]))
    fiber = tg.Material("fiber", diameter=0.037, min_bend_radius=0.050)
    lines = [[[0.1 + 0.08 * i, 0.2 + 0.15 * j, 0.25] for i in range(11)] for j in range(4)]
    recipe.insert(tg.FiberCollection.from_centerlines(lines, fiber, formation_layer=0))
    recipe.needle_layer(0, footprint=tg.RandomFiberFraction(1.0, seed=5), depth=0.05, stiffness=1.0,
                        max_translation=0.01)
    recipe.settle_targets(tolerance=1e-6, max_iterations=200)
    recipe.release_needles()
    result = recipe.run(settings)
    assert result.needling == [(0, 4, 0.05)]
    settled = [e for e in result.events if e.startswith("advance formation targets")]
    assert len(settled) == 1
    # the command walks at most min(0.01, 0.25 d) = 0.00925 per iteration
    assert int(settled[0].split(" in ")[1].split()[0]) >= 6
