"""The layered, needled felt through Recipe.run (formation steps, ply
relaxation, needling of 5 % of the 19 um fibers 30 um through the
thickness, release, final contact and curvature cleanup).

- At a scaled size (4 plies of 200 x 7 um + 27 x 19 um fibers on a 1 mm
  footprint) the whole recipe runs once on the HIP backend and once with
  the oracle standing in for the device world: iterations, events, status
  and every position are equal bit for bit.
- At full size (20 plies x 2,500 fibers, 2 x 2 mm, 3.3 M segments)
  properties: every fiber is placed, every ply is needled 30 um deep,
  positions are finite and inside the z walls, and the
  cleanup leaves the felt at its measured plateau (penetration below
  1e-6 m, curvature ratio below 1.06 after 512 cleanup iterations). The
  felt does not reach the hard gates: its fine fibers settle bent along
  their whole length at the bend limit, where contact- and
  curvature-dominant stages trade one residual for the other."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle
from tangle_amd import _native

pytestmark = pytest.mark.gpu


def test_scaled_layered_felt_recipe_matches_the_oracle(hip_device, monkeypatch):
    recipe, settings = fx.layered_felt_recipe(plies=4, fine=200, coarse=27, footprint=1e-3, final_iterations=128)
    print("[case] layered felt 4 plies, HIP", flush=True)
    gpu = recipe.run(settings)
    monkeypatch.setattr(_native, "World", lambda packed, config: oracle.OracleWorld(packed, config))
    print("[case] layered felt 4 plies, oracle", flush=True)
    ref = recipe.run(settings)
    assert gpu.iterations == ref.iterations
    assert gpu.events == ref.events
    assert gpu.warnings == ref.warnings
    assert gpu.needling == ref.needling
    assert gpu.converged == ref.converged
    assert np.float32(gpu.max_penetration) == np.float32(ref.max_penetration)
    assert np.float32(gpu.max_curvature_ratio) == np.float32(ref.max_curvature_ratio)
    assert np.array_equal(np.asarray(gpu.positions, np.float32).view(np.uint32),
                          np.asarray(ref.positions, np.float32).view(np.uint32))


def test_full_size_layered_felt_properties(hip_device):
    recipe, settings = fx.layered_felt_recipe(plies=20, final_iterations=512)
    print("[case] layered felt 20 plies x 2,500 fibers", flush=True)
    res = recipe.run(settings)
    assert res.fiber_count == 50_000
    pos = np.asarray(res.positions, np.float64).reshape(-1, 3)
    assert np.all(np.isfinite(pos))
    counts = np.asarray(res.vertex_counts, np.int64)
    radius = np.repeat([f.radius for f in recipe._fibers], counts)
    # the walls clamp in f32: allow one f32 ulp at the top of the 6 mm cell
    lo, hi, ulp = 0.0, 20 * 300e-6, float(np.spacing(np.float32(6e-3)))
    assert np.all(pos[:, 2] >= lo + radius - ulp) and np.all(pos[:, 2] <= hi - radius + ulp)
    # one needle command per ply, 30 um through the thickness
    assert [n[0] for n in res.needling] == list(range(20))
    assert all(abs(n[2] - 30e-6) < 1e-12 for n in res.needling)
    assert sum(n[1] for n in res.needling) > 0
    assert res.max_penetration < 1e-6
    assert res.max_curvature_ratio < 1.06
