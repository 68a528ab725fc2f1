"""Compaction kinematics, compaction metrics and contact capture on the HIP
backend against the oracle, through the C ABI (compact_cell,
compaction_metrics and capture_contacts, crates/tangle_relax/src/device/
world.rs:1931-2334; kernels.rs:712-923; crates/tangle_contact/src/device/
contact.rs:9-230), plus checkpoints of a compacted world and per-fiber
curvature limits (world.rs:2182-2191).

Bars: positions, wall reactions (through the face pressures), the 13 metric
values and every captured pair with its coordinates, gap and angle equal
bit for bit; relaxation after a compaction stays bit-exact (the grid follows
the cell); the reference's known answers hold on the device."""

from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

pytestmark = pytest.mark.gpu

KIN = {"rigid_fiber_centers": 0, "moving_walls": 1, "affine_vertices": 2}


def metrics_array(m: _native.CompactionMetrics) -> np.ndarray:
    fp = m.face_pressures
    return np.array([*m.pressure, m.contact_energy, m.stretch_energy, m.bending_energy, m.total_energy,
                     fp[0][0], fp[0][1], fp[1][0], fp[1][1], fp[2][0], fp[2][1]], dtype=np.float32)


def assert_bits(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    diff = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    assert diff.size == 0, f"{what}: {diff.size} values differ, first at {diff[0]}: {a[diff[0]]} vs {b[diff[0]]}"


def assert_capture_equal(gpu, ref):
    gp, gc, gg, ga, go = gpu
    op, oc, og, oa, oo = ref
    assert go == oo
    assert np.array_equal(gp, op), "captured pairs differ"
    assert_bits(gc, oc, "closest-point coordinates")
    assert_bits(gg, og, "surface gaps")
    assert_bits(ga, oa, "crossing angles")


def single_fiber():
    return fx.packed([[[1.0, 1.0, 1.0], [2.0, 1.0, 1.0]]], 0.1, [4.0] * 3, [False] * 3)


@pytest.mark.parametrize("kinematics", list(KIN))
def test_single_fiber_kinematics_known_answers(hip_device, kinematics):
    """world_tests.rs:876-932: rigid keeps length 1 and moves the centre to
    0.75; affine halves the length; the moving wall projects the far vertex."""
    print(f"[case] single fiber {kinematics}", flush=True)
    cfg = fx.config(max_step=0.1)
    upper = [2.0, 4.0, 4.0] if kinematics != "moving_walls" else [1.5, 4.0, 4.0]
    with _native.World(single_fiber(), cfg) as gpu, oracle.OracleWorld(single_fiber(), cfg) as ref:
        gpu.compact_cell([0.0, 0.0, 0.0], upper, kinematics)
        ref.compact_cell([0.0, 0.0, 0.0], upper, KIN[kinematics])
        assert gpu.cell_bounds() == ref.cell_bounds() == ((0.0, 0.0, 0.0), tuple(upper))
        g, o = gpu.download_positions(), ref.download_positions()
        assert_bits(g, o, "positions")
        if kinematics == "rigid_fiber_centers":
            assert abs(g[3] - g[0] - 1.0) < 1e-6 and abs(0.5 * (g[0] + g[3]) - 0.75) < 1e-6
        elif kinematics == "affine_vertices":
            assert abs(g[3] - g[0] - 0.5) < 1e-6
        else:
            assert g[3] == np.float32(np.float32(1.5) - np.float32(0.1))
        m = gpu.compaction_metrics(0.5, _native.CompactionEnergyModel(2.0, 3.0, 1.0))
        assert_bits(metrics_array(m), ref.compaction_metrics(0.5, 2.0, 3.0, 1.0), "metrics")
        if kinematics == "moving_walls":
            assert m.face_pressures[0][1] > 0.0 and m.face_pressures[0][0] == 0.0


@pytest.mark.parametrize("builder,gap,tol,step", [(fx.periodic_pair, 0.1, 1e-5, 0.1),
                                                  (fx.si_crossing, 20e-6, 0.01e-6, 1e-6),
                                                  (fx.unwrapped_images, 20e-6, 0.01e-6, 1e-6)])
def test_capture_reference_cases(hip_device, builder, gap, tol, step):
    """world_tests.rs:934-1026: one candidate each, gap -0.1 or -20 um; after
    the capture the periodic pair relaxes to the tolerance in 40 iterations
    (the lists rebuild on the relaxation grid)."""
    print(f"[case] capture {builder.__name__}", flush=True)
    cfg = fx.config(max_step=step)
    with _native.World(builder(), cfg) as gpu, oracle.OracleWorld(builder(), cfg) as ref:
        cap = gpu.capture_contacts(0.0, 16)
        assert not cap.overflow and len(cap.candidates) == 1
        assert abs(cap.candidates[0].surface_gap + gap) < tol
        assert_capture_equal(gpu.capture_contact_arrays(0.0, 16), ref.capture_contact_arrays(0.0, 16))
        if builder is fx.periodic_pair:
            run = fx.config(penetration_tolerance=1e-5, correction_fraction=0.8, max_step=0.05)
            s, r = gpu.run_batch(run, 40), ref.run_batch(run, 40)
            assert s.total_iterations == r.total_iterations
            assert_bits(gpu.download_positions(), ref.download_positions(), "positions after capture")
            assert s.converged and s.max_penetration <= 1e-5


def test_capture_random_mat_and_overflow(hip_device):
    rng = np.random.default_rng(7)
    fibers = []
    for _ in range(24):
        c = rng.uniform(0.0, 1.0, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        fibers.append([c + (t - 1.5) * 0.12 * d for t in range(4)])
    pk = fx.packed(fibers, np.full(24, 0.03), [1.0, 1.0, 1.0], [True, True, False])
    cfg = fx.config(max_step=0.01)
    print("[case] capture random mat", flush=True)
    with _native.World(pk, cfg) as gpu, oracle.OracleWorld(pk, cfg) as ref:
        full = ref.capture_contact_arrays(0.02, 4096)
        assert_capture_equal(gpu.capture_contact_arrays(0.02, 4096), full)
        assert len(full[0]) > 10
        # overflow: the kept pairs are some of the full list, sorted
        pairs, coords, gaps, angles, overflow = gpu.capture_contact_arrays(0.02, 5)
        assert overflow and len(pairs) == 5
        keys = {tuple(p) for p in full[0].tolist()}
        assert all(tuple(p) in keys for p in pairs.tolist())
        k = pairs[:, 0].astype(np.int64) * 100000 + pairs[:, 1]
        assert np.all(np.diff(k) > 0)


def bimodal_bounded():
    """The small bimodal felt with z bounded, so the walls act."""
    base = fx.bimodal_recipe(scale=0.05)._packed(False)
    return _native.PackedAssembly(base.counts, base.placed, base.intrinsic, base.radius, base.min_bend_radius,
                                  list(base.struct.cell_origin), list(base.struct.cell_lengths),
                                  [True, True, False])


@pytest.mark.parametrize("kinematics", list(KIN))
def test_felt_compaction_then_relaxation_bitwise(hip_device, kinematics):
    """40 iterations, compaction of x and z by 4 % (y unchanged), metrics,
    40 more iterations on the compacted grid, capture: all bit-exact."""
    print(f"[case] bimodal felt compaction {kinematics}", flush=True)
    pk = bimodal_bounded()
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    with _native.World(pk, cfg) as gpu, oracle.OracleWorld(pk, cfg) as ref:
        s, r = gpu.run_batch(cfg, 40), ref.run_batch(cfg, 40)
        assert s.total_iterations == r.total_iterations
        # bounded z: the curvature sweep clamps to the walls on every lane's axis
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions before compaction")
        lo, hi = gpu.cell_bounds()
        span = [hi[k] - lo[k] for k in range(3)]
        new_lo = [lo[0] + 0.02 * span[0], lo[1], lo[2] + 0.02 * span[2]]
        new_hi = [hi[0] - 0.02 * span[0], hi[1], hi[2] - 0.02 * span[2]]
        gpu.compact_cell(new_lo, new_hi, kinematics)
        ref.compact_cell(new_lo, new_hi, KIN[kinematics])
        assert gpu.cell_bounds() == ref.cell_bounds()
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions after compaction")
        model = _native.CompactionEnergyModel(1.0e3, 2.0, 0.5)
        assert_bits(metrics_array(gpu.compaction_metrics(cfg.correction_fraction, model)),
                    ref.compaction_metrics(cfg.correction_fraction, 1.0e3, 2.0, 0.5), "metrics after compaction")
        s, r = gpu.run_batch(cfg, 40), ref.run_batch(cfg, 40)
        assert (s.total_iterations, s.batch_iterations) == (r.total_iterations, r.batch_iterations)
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions after relaxing")
        gm = metrics_array(gpu.compaction_metrics(cfg.correction_fraction, model))
        assert_bits(gm, ref.compaction_metrics(cfg.correction_fraction, 1.0e3, 2.0, 0.5), "metrics after relaxing")
        assert gm[3] > 0.0 and gm[5] >= 0.0
        if kinematics == "moving_walls":
            assert gm[11] > 0.0 or gm[12] > 0.0  # the z walls push back
        gap = 1e-6
        assert_capture_equal(gpu.capture_contact_arrays(gap, 1 << 16), ref.capture_contact_arrays(gap, 1 << 16))
        # the world relaxes on after a capture
        s, r = gpu.run_batch(cfg, 10), ref.run_batch(cfg, 10)
        assert s.total_iterations == r.total_iterations
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions after capture")


def test_compacted_checkpoint_restores_cell(hip_device):
    print("[case] compacted checkpoint", flush=True)
    pk = bimodal_bounded()
    cfg = fx.hard_gate_settings(7e-6, neighbor_capacity=256).to_config(7e-6)
    with _native.World(pk, cfg) as a:
        a.run_batch(cfg, 20)
        lo, hi = a.cell_bounds()
        new_hi = [hi[0], hi[1], hi[2] - 0.05 * (hi[2] - lo[2])]
        a.compact_cell(lo, new_hi, "moving_walls")
        ck = a.checkpoint()
        assert ck.cell_upper == tuple(new_hi) or np.allclose(ck.cell_upper, np.float32(new_hi))
        with _native.World.restore(pk, cfg, ck) as b:
            assert b.cell_bounds() == a.cell_bounds()
            assert_bits(b.download_positions(), a.download_positions(), "restored positions")
            sa, sb = a.run_batch(cfg, 30), b.run_batch(cfg, 30)
            assert sa.total_iterations == sb.total_iterations
            assert_bits(b.download_positions(), a.download_positions(), "continued positions")
            m = _native.CompactionEnergyModel()
            assert_bits(metrics_array(a.compaction_metrics(cfg.correction_fraction, m)),
                        metrics_array(b.compaction_metrics(cfg.correction_fraction, m)), "continued metrics")


def test_fiber_maximum_curvatures_bitwise(hip_device):
    print("[case] fiber curvature limits", flush=True)
    pk = fx.bimodal_recipe(scale=0.05)._packed(False)
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    F = len(pk.counts)
    limits = np.where(np.arange(F) % 3 == 0, 0.0, 1.0 / 150e-6).astype(np.float32)
    with _native.World(pk, cfg) as gpu, oracle.OracleWorld(pk, cfg) as ref:
        gpu.set_fiber_maximum_curvatures(limits)
        ref.set_fiber_maximum_curvatures(limits)
        s, r = gpu.run_batch(cfg, 30), ref.run_batch(cfg, 30)
        assert s.total_iterations == r.total_iterations
        assert_bits(np.float32([s.max_curvature_ratio]), np.float32([r.max_curvature_ratio]), "curvature ratio")
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions")
        with pytest.raises(ValueError):
            gpu.set_fiber_maximum_curvatures(limits[:-1])
        assert not math.isnan(s.max_curvature_ratio)


def test_fit_cell_to_capsules_bitwise(hip_device):
    """The cell fit's path through the world: capsule bounds of the active
    fibers (+ padding) on the bounded axis, moving walls, relaxation on the
    fitted grid: bit-exact against the oracle."""
    from tangle_amd import _active_capsule_bounds

    print("[case] fit cell to capsules", flush=True)
    pk = bimodal_bounded()
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    recipe = fx.bimodal_recipe(scale=0.05)
    starts = np.concatenate([[0], np.cumsum(pk.counts.astype(np.int64))[:-1]])
    with _native.World(pk, cfg) as gpu, oracle.OracleWorld(pk, cfg) as ref:
        gpu.run_batch(cfg, 40)
        ref.run_batch(cfg, 40)
        pos = gpu.download_positions()
        assert_bits(pos, ref.download_positions(), "positions before the fit")
        lo_fit, hi_fit = _active_capsule_bounds(recipe._fibers, pos.reshape(-1, 3), starts,
                                                [True] * len(recipe._fibers), np.float32(1e-6))
        lo, hi = gpu.cell_bounds()
        new_lo = [lo[0], lo[1], max(lo_fit[2], np.float32(lo[2]))]
        new_hi = [hi[0], hi[1], min(hi_fit[2], np.float32(hi[2]))]
        gpu.compact_cell(new_lo, new_hi, "moving_walls")
        ref.compact_cell(new_lo, new_hi, KIN["moving_walls"])
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions after the fit")
        s, r = gpu.run_batch(cfg, 40), ref.run_batch(cfg, 40)
        assert s.total_iterations == r.total_iterations
        assert_bits(gpu.download_positions(), ref.download_positions(), "positions on the fitted cell")


def bounded_bimodal_recipe(scale=0.05):
    import tangle_amd as tg

    side = 1e-3 * scale ** (1.0 / 3.0)
    cell = tg.Cell([side] * 3, periodic="xy")
    recipe = tg.Recipe(cell)
    for spec, name in ((fx.BIMODAL_FINE, "fine"), (fx.BIMODAL_COARSE, "coarse")):
        material = tg.Material(name, spec["diameter"], spec["bend"])
        count = max(1, int(round(spec["count"] * scale)))
        segments = max(2, int(round(spec["segments"] * side / 1e-3)))
        pop = tg.FiberPopulation(material=material, count=count, length=side, segments_per_fiber=segments,
                                 seed=spec["seed"], orientation=tg.SchladitzOrientation(beta=1.0 / 0.3, axis="z"))
        recipe.insert(tg.generate_fiber_population(cell, pop))
    return recipe


def test_recipe_fit_cell_and_bend_radius(hip_device):
    """Recipe.fit_cell_to_active_fibers and Recipe.set_min_bend_radius in a
    run: the events read as the reference writes them, the result's cell is
    the fitted one and every vertex stays inside its walls."""
    import tangle_amd as tg

    print("[case] recipe fit cell and bend radius", flush=True)
    recipe = bounded_bimodal_recipe()
    recipe.relax_for(64)
    recipe.set_min_bend_radius("coarse", 120e-6)
    recipe.fit_cell_to_active_fibers(padding=1e-6)
    recipe.relax_for(64)
    settings = fx.hard_gate_settings(7e-6, max_iterations=2000)
    try:
        res = recipe.run(settings)
    except tg.RecipeError as err:  # the hard gates need not be met in the budget
        pytest.fail(f"recipe failed: {err}")
    bend = [e for e in res.events if e.startswith("set ")]
    fit = [e for e in res.events if e.startswith("fit cell")]
    n_coarse = sum(f.material.name == "coarse" for f in recipe._fibers)
    assert bend == [f'set {n_coarse} "coarse" fibers to minimum bend radius 1.200e-4']
    assert len(fit) == 1 and fit[0].startswith("fit cell to active fibers with 1.000e-6 padding: [")
    cell = res._cell
    side = recipe._cell.lengths[2]
    # the host cell is rebuilt from the f32 bounds (update_host_cell), so
    # the unfitted axes keep their length to f32 precision
    assert abs(cell.lengths[0] - recipe._cell.lengths[0]) < 1e-6 * recipe._cell.lengths[0]
    assert cell.lengths[2] < side
    xyz = res.positions.reshape(-1, 3)
    lo, hi = np.float32(cell.origin[2]), np.float32(cell.origin[2] + cell.lengths[2])
    radius = np.repeat([np.float32(f.radius) for f in recipe._fibers], res.vertex_counts)
    assert np.all(xyz[:, 2] >= lo + radius) and np.all(xyz[:, 2] <= hi + np.float32(1e-9) - radius)
    assert res._bend[[f.material.name for f in recipe._fibers].index("coarse")] == 120e-6
    bad = bounded_bimodal_recipe()
    bad.set_min_bend_radius("wool", 1e-4)
    with pytest.raises(tg.RecipeError, match='unknown material "wool"; known materials: "fine", "coarse"'):
        bad.run(settings)


def lenient_settings(**changes):
    import tangle_amd as tg

    return tg.RelaxationSettings(penetration_tolerance=0.05 * 7e-6, curvature_ratio_tolerance=0.01,
                                 max_iterations=changes.pop("max_iterations", 6000), **changes)


def test_recipe_compact_reaches_volume_fraction(hip_device):
    """Recipe.compact: a bounded felt compacted along z to a nominal volume
    fraction; the step reports grow the fraction, x and y stay, and the
    operation ends TargetReached with the reference's event text."""
    import tangle_amd as tg
    from tangle_amd import compaction as cp

    print("[case] recipe compact to volume fraction", flush=True)
    recipe = bounded_bimodal_recipe()
    recipe.relax_for(128)
    solid = cp.nominal_fiber_volume(recipe._fibers, None)
    side = recipe._cell.lengths
    start_fraction = solid / (side[0] * side[1] * side[2])
    target = round(start_fraction * 1.12, 4)
    spec = tg.CompactionSettings.volume_fraction(target, kinematics="moving_walls", max_curvature_ratio=1.05)
    recipe.compact(spec)
    res = recipe.run(lenient_settings())
    assert len(res.compactions) == 1
    report = res.compactions[0]
    assert report["reason"] == "TargetReached", report
    assert report["nominal_volume_fraction"] >= target * (1 - 1e-3)
    assert report["steps"] >= 2 and report["steps"] == len(res.compaction_steps)
    fractions = [s["nominal_volume_fraction"] for s in res.compaction_steps]
    assert all(b > a for a, b in zip(fractions, fractions[1:]))
    lengths = report["cell_lengths"]
    assert abs(lengths[0] - side[0]) < 1e-6 * side[0] and abs(lengths[1] - side[1]) < 1e-6 * side[1]
    assert lengths[2] < side[2]
    done = [e for e in res.events if e.startswith("compact in ")]
    assert done == [f"compact in {report['steps']} steps to nominal volume fraction "
                    f"{report['nominal_volume_fraction']:.4f} (TargetReached)"]
    # walls moved in on z: every vertex inside the final cell
    cell = res._cell
    xyz = res.positions.reshape(-1, 3)
    radius = np.repeat([np.float32(f.radius) for f in recipe._fibers], res.vertex_counts)
    assert np.all(xyz[:, 2] >= np.float32(cell.origin[2]) + radius)
    assert np.all(xyz[:, 2] <= np.float32(cell.origin[2] + cell.lengths[2]) + np.float32(1e-9) - radius)


def test_recipe_compact_stop_reasons(hip_device):
    """A baseline that cannot relax in its windows stops with the
    reference's guard text; a step whose one-iteration window cannot relax
    rolls the world back to the accepted state and ends Jammed."""
    import tangle_amd as tg

    print("[case] recipe compact stop reasons", flush=True)
    strict = tg.RelaxationSettings(penetration_tolerance=1e-12, curvature_ratio_tolerance=1e-9, max_iterations=400)
    recipe = bounded_bimodal_recipe()
    recipe.compact(tg.CompactionSettings.volume_fraction(0.3, relax_iterations=8, max_relax_windows=2))
    res = recipe.run(strict)
    assert res.compactions[0]["reason"].startswith('GuardReached("initial formation state did not relax '
                                                   '(penalty energy ')
    assert res.compactions[0]["steps"] == 0
    recipe = bounded_bimodal_recipe()
    recipe.relax_for(400)
    recipe.compact(tg.CompactionSettings.volume_fraction(0.3, relax_iterations=1, max_relax_windows=1,
                                                         initial_log_strain=0.05, max_log_strain=0.05,
                                                         kinematics="affine_vertices", max_curvature_ratio=1.05))
    res = recipe.run(lenient_settings(max_iterations=1200))
    report = res.compactions[0]
    assert report["reason"] == "Jammed", report
    # the cell of the accepted (baseline) state was restored
    assert abs(report["cell_lengths"][2] - recipe._cell.lengths[2]) < 1e-6 * recipe._cell.lengths[2]


def test_recipe_junction_capture_matches_oracle(hip_device):
    """Recipe.capture_junctions and relax_and_capture: the first capture's
    junctions equal those the host policy forms from the oracle's capture
    of the same state, and the events read as the reference writes them."""
    import tangle_amd as tg
    from tangle_amd import junctions as jn

    print("[case] recipe junction capture", flush=True)
    recipe = bounded_bimodal_recipe()
    recipe.relax_for(64)
    policy = tg.JunctionPolicy("bonds", max_surface_gap=0.5e-6, max_per_fiber_pair=2)
    recipe.capture_junctions(policy)
    recipe.relax_and_capture(iterations=64, capture_every=32, policy=policy.replace(name="periodic"))
    settings = lenient_settings(max_iterations=600)
    res = recipe.run(settings)
    first = res.junction_captures[0]
    assert first["created"] > 0 and len(res.junction_captures) == 3
    events = [e for e in res.events if "junction" in e]
    assert events[0] == f"capture {first['created']} junctions with 'bonds' from {first['candidates']} candidates"
    assert events[1].startswith("periodic capture created ") and events[-1] == (
        "relax for 64 iterations with junction capture every 32")
    assert res.junction_count == sum(c["created"] for c in res.junction_captures)
    # the same state on the oracle: activation of both steps, 64 iterations
    packed = recipe._packed(False)
    cfg = settings.to_config(min(2.0 * f.radius for f in recipe._fibers))
    cfg.force_full_iterations = 1
    with oracle.OracleWorld(packed, cfg) as ref:
        ref.activate_formation_step(1)
        ref.run_batch(cfg, 64)
        cap = ref.capture_contact_arrays(np.float32(policy.max_surface_gap), policy.candidate_capacity)
    store = []
    intrinsic = np.asarray(packed.intrinsic, dtype=np.float32).reshape(-1, 3)
    report = jn.form_junctions(store, recipe._fibers, packed.counts, intrinsic, cap, policy)
    assert report == first
    assert store == res.junctions[:first["created"]]
    # the OVITO frame carries one bond glyph per junction after the segments
    import tempfile
    from pathlib import Path
    with tempfile.TemporaryDirectory() as tmp:
        dump = Path(tmp) / "felt.dump"
        res.write_ovito(dump, bond_radius_ratio=0.5)
        lines = dump.read_text().splitlines()
    segments = int((res.vertex_counts - 1).sum())
    assert int(lines[3]) == segments + res.junction_count
    assert all(row.split()[2] == "3" for row in lines[-res.junction_count:])
