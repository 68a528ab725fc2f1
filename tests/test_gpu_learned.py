"""Learned warm start on the GPU (no counterpart in the reference).

Checks, through the C ABI:
  * the MFMA lane layout the kernels assume, against a plain matrix product;
  * the segment features against the oracle's float32 restatement, bit-exact;
  * the network's forward pass and its gradients (fused MFMA backward, fixed
    order reduction) against the numpy network with the same bf16 roundings;
  * training on recorded plain-relaxer trajectories lowers the loss;
  * a proposal and its revert restore the positions bit for bit;
  * with a trained model, the same plan reaches the target in fewer
    iterations than without it, on a specimen the model never saw.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import fixtures as fx
from oracle import mlp_ref
import tangle_amd as tg
from tangle_amd import _native, learned

pytestmark = pytest.mark.gpu


def random_parameters(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp_ref.PARAMS, dtype=np.float32)
    p[mlp_ref.OFF_W1:mlp_ref.OFF_B1] = rng.uniform(-0.3, 0.3, 4096)
    p[mlp_ref.OFF_B1:mlp_ref.OFF_W2] = rng.uniform(-0.1, 0.1, 64)
    p[mlp_ref.OFF_W2:mlp_ref.OFF_B2] = rng.uniform(-0.3, 0.3, 4096)
    p[mlp_ref.OFF_B2:mlp_ref.OFF_W3] = rng.uniform(-0.1, 0.1, 64)
    p[mlp_ref.OFF_W3:mlp_ref.OFF_B3] = rng.uniform(-0.3, 0.3, 1024)
    p[mlp_ref.OFF_B3:mlp_ref.PARAMS] = rng.uniform(-0.1, 0.1, 16)
    return p


def random_features(rows, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2.0, 2.0, (rows, 64)).astype(np.float32)
    x[:, 51:] = 0.0
    return mlp_ref.f32_to_bf16_bits(x)


def test_mfma_layout_matches_matrix_product(hip_device):
    rng = np.random.default_rng(7)
    a = mlp_ref.f32_to_bf16_bits(rng.standard_normal((16, 16)).astype(np.float32))
    b = mlp_ref.f32_to_bf16_bits(rng.standard_normal((16, 16)).astype(np.float32))
    d = np.zeros((16, 16), dtype=np.float32)
    _native.check(hip_device.tgh_mfma_probe(_native.ptr(a, C.c_uint16), _native.ptr(b, C.c_uint16),
                                            _native.ptr(d, C.c_float)))
    ref = mlp_ref.bf16_bits_to_f32(a).astype(np.float64) @ mlp_ref.bf16_bits_to_f32(b).astype(np.float64)
    assert np.max(np.abs(d - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("rows", [1, 37, 1000])
def test_forward_matches_numpy_network(hip_device, rows):
    print(f"[case] forward rows={rows}", flush=True)
    params = random_parameters(rows)
    x = random_features(rows, rows + 1)
    model = learned.WarmStartModel()
    model.set_parameters(params)
    y = model.forward(x)
    ref, _ = mlp_ref.forward(params, mlp_ref.bf16_bits_to_f32(x), bf16=True)
    scale = np.max(np.abs(ref))
    # FP32 accumulation against float64; a bf16 rounding of a hidden unit
    # may land on the other side of a tie
    assert np.max(np.abs(y - ref)) <= 2e-3 * scale + 1e-6
    model.close()


@pytest.mark.parametrize("rows", [5, 300, 9000])
def test_gradients_match_numpy_network(hip_device, rows):
    print(f"[case] gradients rows={rows}", flush=True)
    params = random_parameters(100 + rows)
    x = random_features(rows, rows + 2)
    t = np.random.default_rng(rows).uniform(-0.5, 0.5, (rows, 3)).astype(np.float32)
    model = learned.WarmStartModel()
    model.set_parameters(params)
    grad, loss = model.gradients(x, t)
    ref_loss, ref_grad = mlp_ref.loss_and_gradients(params, mlp_ref.bf16_bits_to_f32(x), t, bf16=True)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    blocks = [(mlp_ref.OFF_W1, mlp_ref.OFF_B1), (mlp_ref.OFF_B1, mlp_ref.OFF_W2), (mlp_ref.OFF_W2, mlp_ref.OFF_B2),
              (mlp_ref.OFF_B2, mlp_ref.OFF_W3), (mlp_ref.OFF_W3, mlp_ref.OFF_W3 + 3 * 64),
              (mlp_ref.OFF_B3, mlp_ref.OFF_B3 + 3)]
    for lo, hi in blocks:
        g, r = grad[lo:hi], ref_grad[lo:hi]
        assert np.max(np.abs(g - r)) <= 1e-2 * np.max(np.abs(r)) + 1e-7, (lo, hi)
    # the unused outputs get no gradient
    assert not np.any(grad[mlp_ref.OFF_W3 + 3 * 64:mlp_ref.OFF_B3])
    assert not np.any(grad[mlp_ref.OFF_B3 + 3:mlp_ref.PARAMS])
    model.close()


def _segments(packed):
    counts = packed.counts.astype(np.int64)
    v0, radius = [], []
    start = 0
    for f, n in enumerate(counts):
        for k in range(n - 1):
            v0.append(start + k)
            radius.append(np.float32(packed.radius[f]))
        start += n
    return np.array(v0), np.array(radius, dtype=np.float32)


def _midpoints(positions, v0):
    p = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    return (np.float32(0.5) * (p[v0] + p[v0 + 1])).astype(np.float32)


@pytest.mark.parametrize("case", ["dense_mat", "bimodal", "bimodal_history"])
def test_features_match_oracle(hip_device, case):
    print(f"[case] features {case}", flush=True)
    if case == "dense_mat":
        packed, cfg, iterations = fx.dense_mat(), fx.config(max_step=0.01, force_full_iterations=1), 3
    else:
        recipe = fx.bimodal_recipe(scale=0.01)
        packed = recipe._packed(False)
        cfg = tg.RelaxationSettings().to_config(7e-6)
        cfg.neighbor_capacity = 64
        cfg.force_full_iterations = 1
        iterations = 5
    v0, radius = _segments(packed)
    history = None
    with _native.World(packed, cfg) as world:
        if case == "bimodal_history":
            learned.mark_history(world)
            history = _midpoints(world.download_positions(), v0)
        world.run_batch(cfg, iterations)
        feats = learned.world_features(world)  # refreshes stale lists first
        counts, lists = world.neighbor_lists()
        positions = world.download_positions()
        info = world.info()
    assert int(counts.max()) <= int(info.neighbor_block)
    lengths = packed.struct.cell_lengths
    periodic = [bool(packed.struct.cell_periodic[k]) for k in range(3)]
    ref = mlp_ref.segment_features(positions, v0, radius, lists, counts, periodic,
                                   [np.float32(lengths[k]) for k in range(3)], history=history)
    if history is not None:
        assert np.any(ref[:, 51:55])
        # log10 may differ by an ulp between the device and numpy: compare the
        # motion magnitude feature to bf16 resolution, the rest bit for bit
        got = mlp_ref.bf16_bits_to_f32(feats[:, 54])
        want = mlp_ref.bf16_bits_to_f32(ref[:, 54])
        assert np.max(np.abs(got - want)) <= 2.0 ** -7
        feats, ref = feats.copy(), ref.copy()
        feats[:, 54] = ref[:, 54] = 0
    assert np.array_equal(feats, ref)


def _small_recipe(seed_offset):
    return fx.bimodal_recipe(scale=0.02, seed_offset=seed_offset)


def test_training_lowers_loss_and_revert_restores(hip_device):
    recipe = _small_recipe(0)
    packed = recipe._packed(False)
    cfg = tg.RelaxationSettings().to_config(7e-6)
    model = learned.WarmStartModel(seed=3)
    rows = learned.collect_trajectories(model, packed, cfg, iterations=512, interval=64)
    assert rows == model.dataset_rows and rows > 0
    losses = model.train(epochs=30, batch_rows=2048, learning_rate=1e-3, seed=1)
    print("[losses]", losses[0], losses[-1], flush=True)
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.5 * losses[0]
    with _native.World(packed, cfg) as world:
        learned.propose(world, model, 1.0)  # without a history mark nothing moves
        assert np.array_equal(world.download_positions(), packed.placed.astype(np.float32))
        learned.revert(world)
        learned.mark_history(world)
        world.run_batch(cfg, 16)
        before = world.download_positions()
        learned.propose(world, model, 1.0)
        moved = world.download_positions()
        assert not np.array_equal(before, moved)
        learned.revert(world)
        assert np.array_equal(before, world.download_positions())
        with pytest.raises(ValueError):
            learned.revert(world)
    model.close()


def test_warm_start_reaches_the_target_in_fewer_iterations(hip_device):
    """What the trained network gains: on a bimodal specimen it never saw,
    the same relaxation reaches the same (tight) tolerances in clearly fewer
    iterations with learned proposals than without, and the plain projector
    still judges the final state."""
    d = 7e-6
    train = [fx.bimodal_recipe(scale=0.05, seed_offset=o)._packed(False) for o in (7, 13)]
    model, losses = learned.train_warm_start(train, d, interval=32, default_iterations=1536,
                                             cleanup_iterations=0, epochs=40, batch_rows=4096,
                                             learning_rate=2e-3)
    assert losses[-1] < losses[0]
    results = {}
    for label, ws in (("plain", None), ("learned", learned.WarmStart(model, interval=32, scale=2.0, settle=4,
                                                                     active_iterations=100000))):
        recipe = fx.bimodal_recipe(scale=0.05, seed_offset=101)
        recipe.relax_until_converged(max_iterations=20000)
        settings = tg.RelaxationSettings(penetration_tolerance=1e-3 * d, curvature_ratio_tolerance=1e-4,
                                         max_iterations=20000)
        result = recipe.run(settings, warm_start=ws)
        print(f"[warm start] {label}: {result.iterations} iterations, {result.wall_seconds:.2f}s, "
              f"{result.events[-1]}", flush=True)
        assert result.converged
        assert result.max_penetration <= 1e-3 * d and result.max_curvature_ratio <= 1.0 + 1e-4
        results[label] = result.iterations
    assert results["learned"] <= 0.7 * results["plain"], results
    model.close()
