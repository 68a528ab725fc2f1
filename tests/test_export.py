"""OVITO export (CPU): the reference's known answers for the fiber-segments
glyphs (crates/tangle_export/src/ovito.rs:985-1012, 1060-1100), the dump
layout and Rust's `{:.17e}` number form, checked against values the
mathematics fixes (a circle's curvature, the quaternion of a known turn)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import tangle_amd as tg
from tangle_amd import _Fiber, RunResult
from tangle_amd import export


def test_rust_exponent_form():
    assert export.rust_exp(0.5) == "5.00000000000000000e-1"
    assert export.rust_exp(1.5) == "1.50000000000000000e0"
    assert export.rust_exp(0.0) == "0.00000000000000000e0"
    assert export.rust_exp(-0.0) == "-0.00000000000000000e0"
    assert export.rust_exp(-1234.5) == "-1.23450000000000000e3"
    assert export.rust_exp(7e-6) == "6.99999999999999989e-6"  # the f64 nearest 7e-6, 18 digits
    assert export.rust_exp(float("inf")) == "inf" and export.rust_exp(float("nan")) == "NaN"


def test_straight_fiber_becomes_one_oriented_capsule():
    # ovito.rs:985-1010
    caps = export.fiber_capsules([(7, 0.1, np.array([[0.25, 1.0, 1.0], [1.75, 1.0, 1.0]]),
                                   np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), None)], atom_type=3)
    assert len(caps) == 1
    c = caps[0]
    assert c["fiber_id"] == 7 and c["fiber_points"] == 2
    assert abs(c["length"] - 1.5) < 1e-12 and c["position"] == (1.0, 1.0, 1.0)
    assert c["natural_curvature"] == 0.0 and c["current_curvature"] == 0.0
    assert abs(c["orientation"][1] - math.sqrt(0.5)) < 1e-12 and abs(c["orientation"][3] - math.sqrt(0.5)) < 1e-12


def test_quaternion_turns_z_onto_the_segment():
    for d in ([1.0, 2.0, -0.5], [0.0, 0.0, 3.0], [0.3, -0.2, -1.0]):
        n = math.sqrt(sum(v * v for v in d))
        i, j, k, w = export.z_axis_orientation(d, n)
        # rotate e_z by the quaternion (w, i, j, k)
        ez = (2 * (i * k + w * j), 2 * (j * k - w * i), 1 - 2 * (i * i + j * j))
        assert np.allclose(ez, np.array(d) / n, atol=1e-12)
    assert export.z_axis_orientation([0, 0, -1.0], 1.0) == (1.0, 0.0, 0.0, 0.0)


def test_curvature_of_circle_samples():
    # curvature.rs:91-110: samples of a radius-2 circle have curvature 1/2
    pts = [(2 * math.cos(t), 2 * math.sin(t), 0.0) for t in np.linspace(0, math.pi / 2, 9)]
    k = export.vertex_curvatures(pts)
    assert k[0] == 0.0 and k[-1] == 0.0
    assert all(abs(v - 0.5) < 1e-12 for v in k[1:-1])


def test_dump_layout_and_script(tmp_path):
    cell = tg.Cell([4.0, 4.0, 4.0], periodic="x")
    rest = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [1.0, 0.0, 0.0]])
    placed = np.array([[1.0, 1.0, 1.0], [1.5, 1.25, 1.0], [2.0, 1.0, 1.0]], dtype=np.float32)
    fiber = _Fiber(tg.Material("f", 0.2, 1.0), 0.1, placed.astype(np.float64), rest)
    res = RunResult(1, 42, True, 0.0, 0.0, 0.0, 2, 3, 1, 0, 0, [], [], 0, 0.0, placed.reshape(-1), np.array([3]),
                    _fibers=[fiber], _cell=cell)
    dump, script = tmp_path / "final.dump", tmp_path / "view.py"
    res.write_ovito(dump, view_script_path=script, session_path=tmp_path / "final.ovito",
                    coloring="curvature_ratio")
    lines = dump.read_text().splitlines()
    assert lines[:5] == ["ITEM: TIMESTEP", "42", "ITEM: NUMBER OF ATOMS", "2", "ITEM: BOX BOUNDS pp ff ff"]
    assert lines[5] == "0.00000000000000000e0 4.00000000000000000e0"
    assert lines[8] == export.ATOM_HEADER
    first = lines[9].split()
    assert first[:3] == ["1", "1", "1"] and len(first) == 21 and first[-1] == "0"
    assert float(first[3]) == 0.1 and first[13:15] == ["0", "3"]
    # the middle vertex turns by 2 atan(0.5): both glyphs carry its curvature
    theta = 2 * math.atan(0.5)
    dual = math.sqrt(0.25 + 0.0625)
    kappa = 2 * math.sin(theta / 2) / dual
    assert abs(float(first[16]) - kappa) < 1e-12 and abs(float(lines[10].split()[16]) - kappa) < 1e-12
    assert abs(float(first[17]) - kappa / 1.0) < 1e-12          # ratio against 1 / min_bend_radius = 1
    assert abs(float(first[19]) - 1.0) < 1e-12                  # log2(total / segment) = log2(2)
    text = script.read_text()
    assert "trajectory_path = HERE / 'final.dump'" in text
    assert "ColorCodingModifier(property='curvature_ratio', start_value=0.0, end_value=1.0)" in text
    assert "ParticlesVis.Shape.Spherocylinder" in text and "ovito.scene.save" in text
    with pytest.raises(ValueError):
        res.write_ovito(dump, coloring="rainbow")


def _touching_cross(periodic):
    # bridges.rs:186-228: two fibers of radius 0.1 crossing and touching, one
    # junction (id 9) at rest arc length 0.3 on both
    from tangle_amd.junctions import Junction

    if periodic:
        first, second = [[0.05, 0.2, 0.5], [0.05, 0.8, 0.5]], [[0.85, 0.5, 0.2], [0.85, 0.5, 0.8]]
    else:
        first, second = [[0.2, 0.5, 0.4], [0.8, 0.5, 0.4]], [[0.5, 0.2, 0.6], [0.5, 0.8, 0.6]]
    fibers = [(1, 0.1, np.array(first), np.array(first), None), (2, 0.1, np.array(second), np.array(second), None)]
    return fibers, [Junction(9, "bond", 0, ((1, 0.3), (2, 0.3)))], [periodic, False, False]


def test_bridge_spans_the_crossing_between_centerlines():
    fibers, junctions, periodic = _touching_cross(False)
    bridges = export.junction_bridges(fibers, junctions, [0.0] * 3, [1.0] * 3, periodic, 0.5)
    assert len(bridges) == 1 and bridges[0]["junction_id"] == 9
    assert abs(bridges[0]["radius"] - 0.05) < 1e-12
    assert np.allclose(bridges[0]["start"], [0.5, 0.5, 0.4], atol=1e-12)
    assert np.allclose(bridges[0]["end"], [0.5, 0.5, 0.6], atol=1e-12)


def test_bridge_across_a_periodic_wall_takes_the_nearest_image(tmp_path):
    fibers, junctions, periodic = _touching_cross(True)
    bridges = export.junction_bridges(fibers, junctions, [0.0] * 3, [1.0] * 3, periodic, 0.5)
    b = bridges[0]
    assert abs(np.linalg.norm(np.subtract(b["end"], b["start"])) - 0.2) < 1e-12
    assert np.allclose(b["end"], [-0.15, 0.5, 0.5], atol=1e-12)
    for bad in (0.0, float("nan")):
        with pytest.raises(ValueError):
            export.junction_bridges(fibers, junctions, [0.0] * 3, [1.0] * 3, periodic, bad)
    # the frame carries one bond row of type atom_type + 2 after the capsules
    path = tmp_path / "bonded.dump"
    rows = export.write_ovito_frame(path, fibers, [0.0] * 3, [1.0] * 3, periodic, 0, junctions=junctions,
                                    bond_radius_ratio=0.5)
    text = path.read_text().splitlines()
    assert rows == 3 and text[3] == "3"
    bond = text[-1].split()
    assert bond[:3] == ["3", "0", "3"] and bond[-1] == "9"
    assert bond[3] == "5.000000000e-2" and bond[5] == "2.000000000e-1"
