"""HIP backend against the oracle, through the C ABI (libtangle_hip.so).

Bars (from the task's contract for this path):
  * cell lists and neighbour pair sets bit-exact on identical positions;
  * FP32 positions after a fixed iteration count within 1e-4 * d_min max-abs;
  * the hard admissibility gates agree (the oracle re-measures the GPU result).
The fixture tests also assert the reference's own expectations
(crates/tangle_relax/src/device/world_tests.rs, cited per fixture).
"""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def both(packed, cfg):
    return _native.World(packed, cfg), oracle.OracleWorld(packed, cfg)


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


CASES = fx.REFERENCE_CASES


@pytest.mark.parametrize("name,builder,changes,iterations,expect,d_min", CASES, ids=[c[0] for c in CASES])
def test_reference_fixture(hip_device, name, builder, changes, iterations, expect, d_min):
    print(f"[case] {name} iterations={iterations}", flush=True)
    cfg = fx.config(**changes)
    gpu, ref = both(builder(), cfg)
    with gpu, ref:
        sg = gpu.run_batch(cfg, iterations)
        so = ref.run_batch(cfg, iterations)
        assert expect(sg), sg.as_dict()
        assert expect(so), so.as_dict()
        assert sg.converged == so.converged
        assert sg.total_iterations == so.total_iterations
        assert abs(sg.max_penetration - so.max_penetration) <= 1e-4 * d_min
        assert max_abs(gpu.download_positions(), ref.download_positions()) <= 1e-4 * d_min
        assert gpu.info().neighbor_list_rebuilds == ref.info().neighbor_list_rebuilds


def test_pinned_endpoints_stay_exact(hip_device):
    # world_tests.rs:505-508 and :1184-1186: pinned ends keep their exact f32 values
    cfg = fx.config(penetration_tolerance=1e-5, stretch_stiffness=0.0, bend_stiffness=0.0, constraint_iterations=1,
                    curvature_cleanup_sweeps=1, max_step=0.02)
    with _native.World(fx.zigzag(), cfg) as w:
        w.run_batch(cfg, 30)
        p = w.download_positions()
    assert p[0:3].tolist() == np.float32([0.2, 1.0, 1.0]).tolist()
    assert p[240:243].tolist() == np.float32([1.8, 1.0, 1.0]).tolist()


def test_first_pass_penetration_known_answers(hip_device):
    # world_tests.rs:953-956, 992-995, 1022-1025: surface gaps -0.1, -20e-6, -20e-6
    for builder, gap, tol, step in ((fx.periodic_pair, 0.1, 1e-5, 0.05), (fx.si_crossing, 20e-6, 1e-8, 1e-6),
                                    (fx.unwrapped_images, 20e-6, 1e-8, 1e-6)):
        # a tolerance every state meets stops the batch at its first pass, so
        # the status measures the initial overlap
        cfg = fx.config(max_step=step, penetration_tolerance=1.0)
        with _native.World(builder(), cfg) as w:
            s0 = w.run_batch(cfg, 1)
        assert s0.total_iterations == 0
        assert abs(s0.max_penetration - gap) < tol, (builder.__name__, s0.max_penetration)


def test_self_image_contact(hip_device):
    # world_tests.rs:1029-1064
    cfg = fx.config(penetration_tolerance=1e-4, max_step=0.05)
    with _native.World(fx.self_image(), cfg) as w:
        initial = w.run_batch(cfg, 1)
        assert initial.max_penetration > 0.05
        status = w.run_batch(cfg, 999)
        assert status.max_penetration < 0.005


def test_listed_overflow_and_skin0_paths_agree(hip_device):
    # world_tests.rs:1309-1337
    def relax(**cell):
        cfg = fx.config(penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01, **cell)
        with _native.World(fx.dense_mat(), cfg) as w:
            w.run_batch(cfg, 40)
            return w.download_positions(), w.info().neighbor_list_rebuilds

    every, every_rebuilds = relax(neighbor_skin_scale=0.0)
    listed, listed_rebuilds = relax()
    overflowed, _ = relax(neighbor_capacity=1)
    assert max_abs(listed, every) < 1e-4
    assert max_abs(overflowed, every) < 1e-4
    assert listed_rebuilds < every_rebuilds


def test_bitwise_repeatable(hip_device):
    # world_tests.rs:1487-1510
    for cap in (48, 1):
        cfg = fx.config(penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01, neighbor_capacity=cap)
        runs = []
        for _ in range(2):
            with _native.World(fx.dense_mat(), cfg) as w:
                w.run_batch(cfg, 40)
                runs.append((w.download_positions().view(np.uint32).copy(), w.info().neighbor_list_rebuilds))
        assert runs[0][1] == runs[1][1]
        assert np.array_equal(runs[0][0], runs[1][0])


def _walled_bimodal():
    """The bimodal mix in a closed box: wall clamps and reactions are live."""
    recipe = fx.bimodal_recipe(scale=0.02)
    packed = recipe._packed(False)
    return _native.PackedAssembly(packed.counts, packed.placed, packed.intrinsic, packed.radius,
                                  packed.min_bend_radius, [packed.struct.cell_origin[k] for k in range(3)],
                                  [packed.struct.cell_lengths[k] for k in range(3)], [False] * 3)


@pytest.mark.parametrize("case", ["bimodal", "walled", "zigzag"])
def test_curvature_launch_shapes_agree_bitwise(hip_device, case):
    """The curvature sweeps give the same positions bit for bit: the
    sequential 8-lane walk (triplets carried in registers), the screened
    sweeps (screen, then exact projection on three axis lanes) at 4, 16, 32
    and 64 lanes per fiber, and the one-lane walk in global memory."""
    if case == "zigzag":
        packed = fx.zigzag()
        cfg = fx.config(penetration_tolerance=0.0, stretch_stiffness=0.0, bend_stiffness=0.0, force_full_iterations=1,
                        constraint_iterations=1, curvature_cleanup_sweeps=3, max_step=0.02)
    else:
        packed = fx.bimodal_recipe(scale=0.02)._packed(False) if case == "bimodal" else _walled_bimodal()
        cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    results = {}
    for lanes in (0, 4, 8, 16, 32, 64):
        print(f"[case] {case} curvature_lanes={lanes}", flush=True)
        with _native.World(packed, cfg) as w:
            w.set_tuning(curvature_lanes=lanes)
            assert w.tuning().curvature_lanes == lanes
            w.run_batch(cfg, 40)
            results[lanes] = w.download_positions().view(np.uint32).copy()
    for lanes in (4, 8, 16, 32, 64):
        assert np.array_equal(results[lanes], results[0]), lanes


@pytest.mark.parametrize("capacity", [48, 4])
def test_contact_launch_shapes_agree_bitwise(hip_device, capacity):
    """4, 8, 16 and 32 lanes per segment in the contact pass give the same
    positions bit for bit, on listed, overflow-row and 27-cell paths."""
    packed = fx.bimodal_recipe(scale=0.02)._packed(False)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True, neighbor_capacity=capacity).to_config(7e-6)
    results = {}
    for lanes in (4, 8, 16, 32):
        print(f"[case] contact_lanes={lanes} capacity={capacity}", flush=True)
        with _native.World(packed, cfg) as w:
            w.set_tuning(contact_lanes=lanes)
            w.run_batch(cfg, 30)
            results[lanes] = w.download_positions().view(np.uint32).copy()
    for lanes in (4, 8, 32):
        assert np.array_equal(results[lanes], results[16]), lanes


def test_tuning_rejects_bad_lanes(hip_device):
    with _native.World(fx.zigzag(), fx.config()) as w:
        with pytest.raises(ValueError):
            w.set_tuning(curvature_lanes=12)
        with pytest.raises(ValueError):
            w.set_tuning(contact_lanes=64)


def _lists_equal(gpu, ref):
    gc, go, gs = gpu.cell_list()
    oc, oo, os_ = ref.cell_list()
    assert np.array_equal(gc, oc), "cell counts differ"
    assert np.array_equal(go, oo), "cell offsets differ"
    assert np.array_equal(gs, os_), "cell segment order differs"
    gn, gl = gpu.neighbor_lists()
    on, ol = ref.neighbor_lists()
    assert np.array_equal(gn, on), "neighbour counts differ"
    block = gl.shape[1]
    listed = gn <= block
    for s in np.nonzero(listed)[0]:
        n = gn[s]
        assert np.array_equal(gl[s, :n], ol[s, :n]), f"segment {s} list differs"
    return int(listed.sum()), int((~listed).sum())


@pytest.mark.parametrize("scale", [0.02, 0.1])
def test_lists_bit_exact_bimodal(hip_device, scale):
    recipe = fx.bimodal_recipe(scale=scale)
    packed = recipe._packed(False)
    d_min = 7e-6
    cfg = fx.hard_gate_settings(d_min).to_config(d_min)
    print(f"[case] bimodal scale={scale} vertices={packed.vertex_count}", flush=True)
    gpu, ref = both(packed, cfg)
    with gpu, ref:
        gpu.rebuild_lists()
        ref.rebuild_lists()
        listed, overflowed = _lists_equal(gpu, ref)
        assert listed > 0
        # relax on the GPU, then rebuild both from the GPU's positions
        gpu.run_batch(cfg, 60)
        p = gpu.download_positions()
        gpu.upload_positions(p)
        ref.upload_positions(p)
        gpu.rebuild_lists()
        ref.rebuild_lists()
        _lists_equal(gpu, ref)


def test_lists_bit_exact_overflow_block(hip_device):
    # capacity 4 forces overflowed lists; the global-memory list path (block > 64)
    for cap in (4, 80):
        recipe = fx.bimodal_recipe(scale=0.02)
        packed = recipe._packed(False)
        cfg = fx.hard_gate_settings(7e-6, neighbor_capacity=cap).to_config(7e-6)
        gpu, ref = both(packed, cfg)
        with gpu, ref:
            gpu.rebuild_lists()
            ref.rebuild_lists()
            listed, overflowed = _lists_equal(gpu, ref)
            if cap == 4:
                assert overflowed > 0


def test_config1_positions_match(hip_device):
    recipe = fx.config1_recipe()
    packed = recipe._packed(False)
    cfg = fx.hard_gate_settings(10e-6, force_full_iterations=True).to_config(10e-6)
    gpu, ref = both(packed, cfg)
    with gpu, ref:
        for _ in range(3):
            sg = gpu.run_batch(cfg, 25)
            so = ref.run_batch(cfg, 25)
            assert sg.total_iterations == so.total_iterations
        # no bend limit: no acos anywhere, so the trajectories agree bit for bit
        pg, po = gpu.download_positions(), ref.download_positions()
        assert max_abs(pg, po) <= 1e-4 * 10e-6
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


@pytest.mark.parametrize("scale", [0.02, 0.1])
def test_bimodal_positions_match(hip_device, scale):
    recipe = fx.bimodal_recipe(scale=scale)
    packed = recipe._packed(False)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    gpu, ref = both(packed, cfg)
    with gpu, ref:
        sg = gpu.run_batch(cfg, 50)
        so = ref.run_batch(cfg, 50)
        assert sg.total_iterations == so.total_iterations == 50
        pg, po = gpu.download_positions(), ref.download_positions()
        assert max_abs(pg, po) <= 1e-4 * 7e-6
        # the curvature sweep's acosf is the host libm's algorithm: bit for bit
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


@pytest.mark.parametrize("scale", [0.05])
def test_bimodal_reaches_hard_admissibility(hip_device, scale):
    """A bimodal specimen (same fill, fiber mix and bend limits; box scaled)
    relaxes to the hard gates, and the oracle, measuring the GPU's final
    positions, agrees that every gate holds."""
    recipe = fx.bimodal_recipe(scale=scale)
    settings = fx.hard_gate_settings(7e-6, max_iterations=60000)
    recipe.relax_until_converged(max_iterations=60000)
    result = recipe.run(settings)
    print(f"[bimodal] scale={scale} iterations={result.iterations} wall={result.wall_seconds:.2f}s "
          f"pen={result.max_penetration:.3e} curv={result.max_curvature_ratio:.7f}", flush=True)
    assert result.converged
    assert result.max_penetration <= 1e-3 * 7e-6
    assert result.max_curvature_ratio <= 1.0 + 1e-5
    packed = recipe._packed(False)
    cfg = settings.to_config(7e-6)
    with oracle.OracleWorld(packed, cfg) as ref:
        ref.upload_positions(result.positions)
        check = ref.run_batch(cfg, 1)
        assert check.total_iterations == 0 and check.converged, check.as_dict()
        assert check.max_penetration <= 1e-3 * 7e-6
        assert check.max_curvature_ratio <= 1.0 + 1e-5


def test_create_packed_equals_create(hip_device):
    """DeviceWorld::new from PackedAssembly's f32 arrays
    (tgh_world_create_packed) gives the world tgh_world_create builds from
    the f64 assembly: same sizing, lists and trajectory bit for bit."""
    packed = fx.bimodal_recipe(scale=0.02)._packed(False)
    layout = _native.PackedLayout.from_assembly(packed)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    with _native.World(packed, cfg) as a, _native.World.from_packed(layout, cfg) as b:
        ia, ib = a.info(), b.info()
        assert list(ia.cells) == list(ib.cells) and ia.cell_size == ib.cell_size
        sa, sb = a.run_batch(cfg, 30), b.run_batch(cfg, 30)
        assert sa.as_dict() == sb.as_dict()
        assert np.array_equal(a.download_positions().view(np.uint32), b.download_positions().view(np.uint32))


@pytest.mark.parametrize("walled", [False, True])
@pytest.mark.parametrize("passes", [1, 2, 3, 8, 33])
def test_jacobi_pass_counts_match_bitwise(hip_device, passes, walled):
    """Every count of Jacobi passes per iteration matches the oracle bit for
    bit: one pass alone, two to 32 passes in one launch (the workgroup
    recomputes a halo of two vertices per remaining pass), more than that
    one launch per pass; in a periodic cell and in one walled in z, where
    the passes' wall reactions feed the clamps."""
    if walled:
        from test_gpu_compaction import bounded_bimodal_recipe

        recipe = bounded_bimodal_recipe(scale=0.02)
    else:
        recipe = fx.bimodal_recipe(scale=0.02)
    packed = recipe._packed(False)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    cfg.constraint_iterations = passes
    print(f"[case] passes={passes} walled={walled}", flush=True)
    gpu, ref = both(packed, cfg)
    with gpu, ref:
        sg = gpu.run_batch(cfg, 20)
        so = ref.run_batch(cfg, 20)
        assert sg.total_iterations == so.total_iterations == 20
        assert np.float32(sg.max_displacement) == np.float32(so.max_displacement)
        pg, po = gpu.download_positions(), ref.download_positions()
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


@pytest.mark.parametrize("walled", [False, True])
def test_lists_bit_exact_with_fibers_images_away(hip_device, walled):
    """The list build leaves out stencil cells that no candidate within
    reach can occupy; the cells and pair lists still equal the oracle's when
    fibers sit periodic images away from the primary cell (shifted by
    whole cell lengths, fiber by fiber), on a grid of eight cells a side,
    periodic in all axes or walled in z."""
    if walled:
        from test_gpu_compaction import bounded_bimodal_recipe

        recipe = bounded_bimodal_recipe(scale=0.3)
    else:
        recipe = fx.bimodal_recipe(scale=0.3)
    packed = recipe._packed(False)
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    lengths = np.array([packed.struct.cell_lengths[k] for k in range(3)])
    periodic = np.array([bool(packed.struct.cell_periodic[k]) for k in range(3)])
    counts = np.asarray(packed.counts, dtype=np.int64)
    gpu, ref = both(packed, cfg)
    with gpu, ref:
        print(f"[case] walled={walled} cells={list(gpu.info().cells)}", flush=True)
        assert min(gpu.info().cells) >= 5
        p = gpu.download_positions().astype(np.float64).reshape(-1, 3)
        start = 0
        for f, n in enumerate(counts):
            shift = np.array([(f % 5) - 2, (3 * f % 5) - 2, (f % 3) - 1], dtype=np.float64) * lengths * periodic
            p[start:start + n, :3] += shift
            start += n
        p = p.astype(np.float32)
        for _ in range(2):
            gpu.upload_positions(p)
            ref.upload_positions(p)
            gpu.rebuild_lists()
            ref.rebuild_lists()
            listed, _ = _lists_equal(gpu, ref)
            assert listed > 0
            gpu.run_batch(cfg, 30)
            p = gpu.download_positions()
