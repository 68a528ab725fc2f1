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
              origin=(0.0, 0.0, 0.0), max_attempts=128):
    py = dict(count=count, segments=segments, seed=seed, length=length, radius=radius, amplitude=amplitude,
              orientation=orientation, vector=list(vector), angle=angle, beta=beta,
              min_bend_radius=min_bend_radius, max_attempts=max_attempts, origin=list(origin),
              lengths=list(lengths), periodic=list(periodic))
    c = _native.PopulationSpec()
    c.count, c.segments_per_fiber, c.seed = count, segments, seed
    c.length_min, c.length_max = length
    c.radius_min, c.radius_max = radius
    c.amplitude_min, c.amplitude_max = amplitude
    c.orientation = {"isotropic": 0, "planar": 1, "aligned": 2, "schladitz": 3}[orientation]
    c.orientation_vector[:] = list(vector)
    c.orientation_angle = angle
    c.orientation_beta = beta
    c.min_bend_radius = min_bend_radius or 0.0
    c.max_attempts_per_fiber = max_attempts
    c.cell_origin[:] = list(origin)
    c.cell_lengths[:] = list(lengths)
    c.cell_periodic[:] = [1 if p else 0 for p in periodic]
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
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cpp_generator_matches_restatement(name):
    py, c = spec_both(**CASES[name])
    ri, rp, rr = ref.generate(py)
    ci, cp, cr = _native.generate_population(c)
    assert np.array_equal(np.asarray(ri), ci)
    assert np.array_equal(np.asarray(rp), cp)
    assert np.array_equal(np.asarray(rr), cr)


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
    intrinsic, placed, radius = _native.generate_population(c)
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
    intrinsic, _, _ = _native.generate_population(c)
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
    intrinsic, placed, _ = _native.generate_population(c)
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
