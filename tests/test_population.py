"""Seeded population generator (population.rs:246-813): the product's C++
generator against the pure-Python restatement, bit for bit, plus the
reference's own generator tests and published SplitMix64 values."""

from __future__ import annotations

import math

import numpy as np
import pytest

import tangle_amd as tg
from oracle import population_ref as ref
from tangle_amd import _native


def spec_both(count, segments, seed, length, radius, amplitude, orientation="isotropic", vector=(0, 0, 1),
              angle=0.0, beta=1.0, min_bend_radius=None, lengths=(1.0, 1.0, 1.0), periodic=(True, True, True),
              origin=(0.0, 0.0, 0.0), max_attempts=128, position="uniform", position_axis=2, layers=1, jitter=0.0,
              exponent=1.0, toward_high=True, primary=0.4, cross=0.4, deviation=0.0, layer_seed=1):
    py = dict(count=count, segments=segments, seed=seed, length=length, radius=radius, amplitude=amplitude,
              orientation=orientation, vector=list(vector), angle=angle, beta=beta,
              min_bend_radius=min_bend_radius, max_attempts=max_attempts, origin=list(origin),
              lengths=list(lengths), periodic=list(periodic), position=position, position_axis=position_axis,
              layers=layers, jitter=jitter, exponent=exponent, toward_high=toward_high, primary=primary,
              cross=cross, deviation=deviation, layer_seed=layer_seed)
    c = _native.PopulationSpec()
    c.count, c.segments_per_fiber, c.seed = count, segments, seed
    c.length_min, c.length_max = length
    c.radius_min, c.radius_max = radius
    c.amplitude_min, c.amplitude_max = amplitude
    c.orientation = {"isotropic": 0, "planar": 1, "aligned": 2, "schladitz": 3, "layered_biaxial": 4}[orientation]
    c.orientation_vector[:] = list(vector)
    c.orientation_angle = angle
    c.orientation_beta = beta
    c.min_bend_radius = min_bend_radius or 0.0
    c.max_attempts_per_fiber = max_attempts
    c.cell_origin[:] = list(origin)
    c.cell_lengths[:] = list(lengths)
    c.cell_periodic[:] = [1 if p else 0 for p in periodic]
    c.position = {"uniform": 0, "layered": 1, "gradient": 2}[position]
    c.position_axis, c.position_layers, c.position_toward_high = position_axis, layers, 1 if toward_high else 0
    c.position_jitter_fraction, c.position_exponent = jitter, exponent
    c.orientation_primary_fraction, c.orientation_cross_fraction = primary, cross
    c.orientation_in_plane_deviation, c.orientation_layer_seed = deviation, layer_seed
    return py, c


CASES = {
    "isotropic_wavy_periodic": dict(count=25, segments=8, seed=1, length=(0.4, 0.6), radius=(0.01, 0.015),
                                    amplitude=(0.0, 0.02), min_bend_radius=0.08),
    "bounded_box": dict(count=20, segments=6, seed=7, length=(0.2, 0.3), radius=(0.01, 0.01),
                        amplitude=(0.0, 0.03), lengths=(1.0, 0.8, 0.6), periodic=(False, False, False),
                        min_bend_radius=0.1),
    "planar_tilt": dict(count=15, segments=10, seed=11, length=(0.3, 0.3), radius=(0.005, 0.005),
                        amplitude=(0.0, 0.0), orientation="planar", vector=(0, 0, 1), angle=0.2,
                        periodic=(True, True, False)),
    "aligned_cone": dict(count=15, segments=5, seed=12, length=(0.2, 0.25), radius=(0.01, 0.02),
                         amplitude=(0.001, 0.002), orientation="aligned", vector=(1, 1, 0), angle=0.3),
    "schladitz_planar": dict(count=30, segments=12, seed=20260921, length=(1e-3, 1e-3), radius=(3.5e-6, 3.5e-6),
                             amplitude=(0.0, 5.6e-6), orientation="schladitz", beta=1 / 0.3,
                             lengths=(1e-3, 1e-3, 1e-3), min_bend_radius=200e-6),
    # the felt example's plies (examples/felt_20ply_control_vs_needled/src/recipe.rs:364-396), scaled
    "layered_biaxial_felt": dict(count=40, segments=8, seed=20260921, length=(0.6e-3, 0.8e-3),
                                 radius=(3.5e-6, 3.5e-6), amplitude=(0.0, 2e-6), orientation="layered_biaxial",
                                 vector=(0, 0, 1), angle=math.pi / 1800, primary=0.4, cross=0.4,
                                 deviation=math.pi / 18, layer_seed=20260920, position="layered", layers=5,
                                 jitter=0.2, lengths=(1e-3, 1e-3, 0.5e-3), periodic=(True, True, False),
                                 min_bend_radius=200e-6),
    "gradient_low": dict(count=20, segments=4, seed=5, length=(0.2, 0.3), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                         position="gradient", position_axis=1, exponent=2.5, toward_high=False,
                         periodic=(True, False, True)),
    "gradient_high": dict(count=20, segments=4, seed=6, length=(0.2, 0.3), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                          position="gradient", position_axis=2, exponent=0.5, toward_high=True,
                          periodic=(True, True, False)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cpp_generator_matches_restatement(name):
    py, c = spec_both(**CASES[name])
    ri, rp, rr, rl = ref.generate(py)
    ci, cp, cr, cl = _native.generate_population(c)
    assert np.array_equal(np.asarray(ri), ci)
    assert np.array_equal(np.asarray(rp), cp)
    assert np.array_equal(np.asarray(rr), cr)
    assert [0xFFFFFFFF if v is None else v for v in rl] == cl.tolist()


def test_splitmix64_published_values():
    # splitmix64 (Steele, Lea & Flood; Vigna's reference C code) from state 0
    # yields 0xe220a8397b1dcdaf first.
    r = ref.SplitMix64(0)
    assert r.next_u64() == 0xE220A8397B1DCDAF
    # unit() keeps the top 53 bits
    r = ref.SplitMix64(0)
    assert ref.SplitMix64(0).unit() == (0xE220A8397B1DCDAF >> 11) / float(1 << 53)


def test_same_seed_reproduces_and_seeds_differ():
    # population.rs:820-832
    _, c = spec_both(**CASES["isotropic_wavy_periodic"])
    a = _native.generate_population(c)
    b = _native.generate_population(c)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    c.seed = 2
    d = _native.generate_population(c)
    assert not np.array_equal(a[1], d[1])


def test_bounded_fibers_are_contained():
    # population.rs:834-861: swept fibers stay inside bounded axes
    case = CASES["bounded_box"]
    _, c = spec_both(**case)
    intrinsic, placed, radius, _ = _native.generate_population(c)
    nv = case["segments"] + 1
    for f in range(case["count"]):
        pts = placed[f * nv:(f + 1) * nv]
        r = radius[f]
        for k in range(3):
            assert pts[:, k].min() >= -1e-12 + r
            assert pts[:, k].max() <= case["lengths"][k] - r + 1e-12


def test_intrinsic_shape_respects_bend_limit():
    case = CASES["isotropic_wavy_periodic"]
    py, c = spec_both(**case)
    intrinsic, _, _, _ = _native.generate_population(c)
    nv = case["segments"] + 1
    for f in range(case["count"]):
        kmax = ref.max_polyline_curvature(intrinsic[f * nv:(f + 1) * nv].tolist())
        assert kmax <= (1.0 / case["min_bend_radius"]) * (1 + 1e-10)


def test_polyline_curvature_of_circle_samples():
    # curvature.rs:91-110: samples of a circle of radius 2 have curvature 0.5
    for points, tol in ((9, 1e-12), (33, 1e-11)):
        pts = [[2.0 * math.cos(i * math.pi / 2 / (points - 1)), 2.0 * math.sin(i * math.pi / 2 / (points - 1)), 0.0]
               for i in range(points)]
        assert abs(ref.max_polyline_curvature(pts) - 0.5) < tol
    assert ref.max_polyline_curvature([[0, 0, 0], [0.5, 0, 0], [1, 0, 0]]) == 0.0


def test_schladitz_sampler_concentrates_in_plane():
    # beta > 1 (the reference's fit convention, shape.rs:26-30) lays fibers
    # in the plane normal to the axis: the median |cos| to z of the model is
    # the quantile at 0.5, u = 0.5 / sqrt(beta^2 - (beta^2 - 1) / 4)
    beta = 1 / 0.3
    _, c = spec_both(count=2000, segments=2, seed=3, length=(0.01, 0.01), radius=(0.001, 0.001),
                     amplitude=(0.0, 0.0), orientation="schladitz", beta=beta)
    intrinsic, placed, _, _ = _native.generate_population(c)
    d = placed[2::3] - placed[0::3]
    cos = np.abs(d[:, 2]) / np.linalg.norm(d, axis=1)
    expected = 0.5 / math.sqrt(beta * beta - (beta * beta - 1) * 0.25)
    assert abs(np.median(cos) - expected) < 0.02
    assert abs(expected - 0.171) < 0.001


def test_facade_population_defaults_follow_diameter():
    # generators.rs:418-420, 567-600: unset length 16-24 d, amplitude 0-0.8 d
    fiber = tg.Material("fiber", diameter=10 * tg.um)
    pop = tg.FiberPopulation(material=fiber, count=20)
    cell = tg.Cell([400 * tg.um] * 3, periodic="xyz")
    col = tg.generate_fiber_population(cell, pop)
    assert len(col) == 20
    for line in col.centerlines():
        length = sum(math.dist(a, b) for a, b in zip(line, line[1:]))
        assert length < 25 * 10 * tg.um


def test_layered_fibers_are_labelled_and_contained():
    # population.rs:833-857: 24 layered fibers (4 layers, jitter 0.2) in a
    # bounded unit cube carry a layer below 4 and stay inside the box
    cell = tg.Cell([1.0] * 3, periodic="")
    pop = tg.FiberPopulation(count=24, position=tg.LayeredPosition(4, axis="z", jitter_fraction=0.2))
    col = tg.generate_fiber_population(cell, pop)
    for f in col._fibers:
        assert f.formation_layer is not None and 0 <= f.formation_layer < 4
        assert np.all(f.placed >= f.radius - 1e-12) and np.all(f.placed <= 1.0 - f.radius + 1e-12)
    assert col.formation_layers() == sorted({f.formation_layer for f in col._fibers})


def test_layered_centres_sit_on_their_layer_band():
    # sample_contained_center (population.rs:696-708): the centre along the
    # axis lies within jitter/2 layer spacings of the layer's mid-plane
    py, c = spec_both(count=40, segments=2, seed=4, length=(0.1, 0.1), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                      orientation="planar", vector=(0, 0, 1), position="layered", layers=5, jitter=0.3,
                      periodic=(True, True, False))
    _, placed, radius, layers = _native.generate_population(c)
    lower, upper = 0.01, 0.99  # planar fibers: the z offsets vanish
    spacing = (upper - lower) / 5
    centres = placed[1::3, 2]
    mid = lower + (layers + 0.5) * spacing
    assert np.all(np.abs(centres - mid) <= 0.15 * spacing + 1e-12)


def _undirected(a, b):
    return abs(math.fmod(math.fmod(a - b + 0.5 * math.pi, math.pi) + math.pi, math.pi) - 0.5 * math.pi)


def test_layered_biaxial_mixture():
    # population.rs:919-950: with no deviation or tilt, 40 % of the
    # directions lie on the layer's reference angle and 40 % a quarter turn
    # away (3700-4300 of 10000 each)
    seed, layer = 83, 3
    reference = ref._layer_reference_angle(seed, layer)
    rng = ref.SplitMix64(19)
    spec = dict(orientation="layered_biaxial", vector=[0.0, 0.0, 1.0], primary=0.4, cross=0.4, deviation=0.0,
                angle=0.0, layer_seed=seed)
    primary = cross = 0
    for _ in range(10_000):
        d = ref._direction(spec, rng, layer)
        azimuth = math.atan2(-d[0], d[1])
        if _undirected(azimuth, reference) < 1e-10:
            primary += 1
        elif _undirected(azimuth, reference + 0.5 * math.pi) < 1e-10:
            cross += 1
    assert 3700 <= primary <= 4300 and 3700 <= cross <= 4300
    # the C++ generator draws the same mixture per layer: straight fibers
    # (amplitude 0) point along their sampled direction
    _, c = spec_both(count=3000, segments=2, seed=19, length=(0.1, 0.1), radius=(0.001, 0.001), amplitude=(0.0, 0.0),
                     orientation="layered_biaxial", vector=(0, 0, 1), primary=0.4, cross=0.4, deviation=0.0,
                     angle=0.0, layer_seed=seed, position="layered", layers=4, jitter=0.0)
    _, placed, _, layers = _native.generate_population(c)
    d = placed[2::3] - placed[0::3]
    az = np.arctan2(-d[:, 0], d[:, 1])
    for lay in range(4):
        r = ref._layer_reference_angle(seed, lay)
        sel = az[layers == lay]
        p = sum(_undirected(a, r) < 1e-9 for a in sel) / len(sel)
        q = sum(_undirected(a, r + 0.5 * math.pi) < 1e-9 for a in sel) / len(sel)
        assert 0.33 < p < 0.47 and 0.33 < q < 0.47


def test_generator_rejects_invalid_layered_specs():
    _, c = spec_both(count=2, segments=2, seed=1, length=(0.1, 0.1), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                     orientation="layered_biaxial", position="uniform")
    with pytest.raises(ValueError):
        _native.generate_population(c)
    _, c = spec_both(count=2, segments=2, seed=1, length=(0.1, 0.1), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                     position="layered", layers=0)
    with pytest.raises(ValueError):
        _native.generate_population(c)
    _, c = spec_both(count=2, segments=2, seed=1, length=(0.1, 0.1), radius=(0.01, 0.01), amplitude=(0.0, 0.0),
                     orientation="layered_biaxial", position="layered", layers=2, primary=0.7, cross=0.5)
    with pytest.raises(ValueError):
        _native.generate_population(c)
    with pytest.raises(ValueError):
        tg.FiberPopulation(orientation=tg.LayeredBiaxialOrientation())._spec(tg.Cell([1.0] * 3))
