"""OVITO export of relaxed fibers (crates/tangle_export/src/ovito.rs): the
fiber-segments view, one spherocylinder glyph per segment in the LAMMPS
custom dump format, and the OVITO Python script that loads and styles it.

The numbers are written as the reference writes them: f64 values in Rust's
`{:.17e}` form (17 digits after the point, exponent without '+' or leading
zeros), geometry in f64 from the relaxed f32 positions, curvature by the
discrete definition 2 sin(theta / 2) / dual length
(crates/tangle_core/src/curvature.rs:20-47).
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np

COLORINGS = ("fiber", "curvature_ratio", "refinement_level")
DEFAULT_BOND_RADIUS_RATIO = 1.0  # crates/tangle_export/src/bridges.rs:19
ATOM_HEADER = ("ITEM: ATOMS id mol type AsphericalShape.X AsphericalShape.Y AsphericalShape.Z quati quatj quatk "
               "quatw x y z segment fiber_points natural_curvature current_curvature curvature_ratio "
               "curvature_excess refinement_level junction")


def rust_exp(value: float, precision: int = 17) -> str:
    """Rust's `{:.Ne}` (LowerExp) for an f64: correctly rounded mantissa,
    exponent as a plain signed integer ("e-1", "e0", "e12")."""
    value = float(value)
    if math.isnan(value):
        return "NaN"
    if math.isinf(value):
        return "inf" if value > 0 else "-inf"
    mantissa, exponent = format(value, f".{precision}e").split("e")
    return f"{mantissa}e{int(exponent)}"


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _norm(a):
    return math.sqrt(_dot(a, a))


def vertex_curvatures(points) -> list[float]:
    """measure_vertex_curvature at every internal vertex, 0 at the ends and
    where a segment is degenerate (curvature.rs:20-47, ovito.rs:683-691)."""
    eps = 2.220446049250313e-16
    out = [0.0] * len(points)
    for i in range(len(points) - 2):
        a, b, c = points[i], points[i + 1], points[i + 2]
        incoming, outgoing = _sub(b, a), _sub(c, b)
        li, lo = _norm(incoming), _norm(outgoing)
        if li <= eps or lo <= eps:
            continue
        cosine = min(max(_dot(incoming, outgoing) / (li * lo), -1.0), 1.0)
        angle = math.acos(cosine)
        dual = 0.5 * (li + lo)
        out[i + 1] = 2.0 * math.sin(0.5 * angle) / dual
    return out


def z_axis_orientation(direction, length: float):
    """Unit quaternion (i, j, k, w) turning +z onto `direction`
    (ovito.rs:693-712)."""
    if length <= 1.0e-15:
        return (0.0, 0.0, 0.0, 1.0)
    r = 1.0 / length
    unit = (direction[0] * r, direction[1] * r, direction[2] * r)
    if unit[2] < -1.0 + 1.0e-12:
        return (1.0, 0.0, 0.0, 0.0)
    q = [-unit[1], unit[0], 0.0, 1.0 + unit[2]]
    total = 0.0
    for v in q:
        total += v * v
    magnitude = math.sqrt(total)
    return tuple(v / magnitude for v in q)


def fiber_capsules(fibers, atom_type: int = 1):
    """One glyph per segment of round fibers (build_fiber_capsules,
    ovito.rs:492-575). `fibers`: (fiber id, radius, placed (n, 3) f64,
    intrinsic (n, 3) f64, minimum bend radius or None)."""
    out = []
    for fiber_id, radius, placed, intrinsic, bend_radius in fibers:
        pts = [tuple(float(v) for v in p) for p in placed]
        rest = [tuple(float(v) for v in p) for p in intrinsic]
        natural = vertex_curvatures(rest)
        current = vertex_curvatures(pts)
        total_rest = 0.0
        for a, b in zip(rest, rest[1:]):
            total_rest += _norm(_sub(b, a))
        kmax = None if bend_radius is None else 1.0 / bend_radius
        for j in range(len(pts) - 1):
            nat = max(natural[j], natural[j + 1])
            cur = max(current[j], current[j + 1])
            ratio = 0.0 if kmax is None else cur / kmax
            excess = 0.0 if kmax is None else max(cur - kmax, 0.0)
            seg_rest = _norm(_sub(rest[j + 1], rest[j]))
            level = max(math.log2(total_rest / seg_rest), 0.0) if seg_rest > 0.0 else 0.0
            delta = _sub(pts[j + 1], pts[j])
            length = _norm(delta)
            position = tuple((pts[j][k] + pts[j + 1][k]) * 0.5 for k in range(3))
            out.append(dict(id=len(out) + 1, fiber_id=fiber_id, atom_type=atom_type, radius=float(radius),
                            length=length, orientation=z_axis_orientation(delta, length), position=position,
                            local_segment=j, fiber_points=len(pts), natural_curvature=nat, current_curvature=cur,
                            curvature_ratio=ratio, curvature_excess=excess, refinement_level=level))
    return out


def _round_half_away(x: float) -> float:
    return math.floor(x + 0.5) if x >= 0.0 else -math.floor(-x + 0.5)


def anchor_point(placed, intrinsic, rest_arc_length: float):
    """Placed centerline point of a material point given by its rest arc
    length: walk the rest centerline to it, then take the same segment
    coordinate on the placed one (anchor_point, bridges.rs:115-170)."""
    traversed, last = 0.0, None
    for j in range(len(intrinsic) - 1):
        length = _norm(_sub(intrinsic[j + 1], intrinsic[j]))
        if length <= 2.220446049250313e-16:
            continue
        last = j
        if rest_arc_length <= traversed + length:
            t = min(max((rest_arc_length - traversed) / length, 0.0), 1.0)
            a, b = placed[j], placed[j + 1]
            return tuple(a[k] + (b[k] - a[k]) * t for k in range(3))
        traversed += length
    if last is None:
        raise ValueError("an anchor's fiber has no segment of positive length")
    return tuple(placed[last + 1])


def junction_bridges(fibers, junctions, lower, upper, periodic, radius_ratio: float):
    """One binder bridge per junction, from its first anchor's placed point to
    the other's nearest periodic image, radius `radius_ratio` times the
    thinner fiber's (junction_bridges, bridges.rs:55-110). `fibers` as for
    fiber_capsules; anchors name fibers by id."""
    if not (math.isfinite(radius_ratio) and radius_ratio > 0.0):
        raise ValueError("bond radius ratio must be positive and finite")
    by_id = {f[0]: f for f in fibers}
    lengths = [float(upper[k]) - float(lower[k]) for k in range(3)]
    out = []
    for junction in junctions:
        (f0, s0), *others = junction.anchors
        a = by_id[f0]
        start = anchor_point([tuple(float(v) for v in p) for p in a[2]], [tuple(float(v) for v in p) for p in a[3]],
                             s0)
        for f1, s1 in others:
            b = by_id[f1]
            end = list(anchor_point([tuple(float(v) for v in p) for p in b[2]],
                                    [tuple(float(v) for v in p) for p in b[3]], s1))
            for k in range(3):
                if periodic[k]:
                    end[k] -= _round_half_away((end[k] - start[k]) / lengths[k]) * lengths[k]
            out.append(dict(junction_id=junction.id, start=start, end=tuple(end),
                            radius=radius_ratio * min(float(a[1]), float(b[1]))))
    return out


def write_ovito_frame(path, fibers, lower, upper, periodic, timestep: int, *, atom_type: int = 1,
                      append: bool = False, junctions=(), bond_radius_ratio=None) -> int:
    """Writes (or appends) one fiber-segments frame (write_fiber_segment_frame
    and write_frame_header, ovito.rs:385-440, 747-770), with one binder-bond
    glyph of type atom_type + 2 per junction bridge when `bond_radius_ratio`
    is given (write_bond_rows, ovito.rs:456-490); returns the number of
    glyph rows."""
    capsules = fiber_capsules(fibers, atom_type)
    bridges = ([] if bond_radius_ratio is None or not junctions else
               junction_bridges(fibers, junctions, lower, upper, periodic, float(bond_radius_ratio)))
    f = rust_exp
    lines = ["ITEM: TIMESTEP", str(int(timestep)), "ITEM: NUMBER OF ATOMS", str(len(capsules) + len(bridges)),
             "ITEM: BOX BOUNDS " + " ".join("pp" if p else "ff" for p in periodic)]
    lines += [f"{f(lower[k])} {f(upper[k])}" for k in range(3)]
    lines.append(ATOM_HEADER)
    for c in capsules:
        q, p = c["orientation"], c["position"]
        lines.append(" ".join([
            str(c["id"]), str(c["fiber_id"]), str(c["atom_type"]), f(c["radius"]), f(c["radius"]), f(c["length"]),
            f(q[0]), f(q[1]), f(q[2]), f(q[3]), f(p[0]), f(p[1]), f(p[2]), str(c["local_segment"]),
            str(c["fiber_points"]), f(c["natural_curvature"]), f(c["current_curvature"]),
            f(c["curvature_ratio"]), f(c["curvature_excess"]), f(c["refinement_level"]), "0"]))
    g = lambda v: rust_exp(v, 9)  # noqa: E731  (write_fiber_glyph's {:.9e})
    for i, bridge in enumerate(bridges):
        delta = _sub(bridge["end"], bridge["start"])
        length = _norm(delta)
        q = z_axis_orientation(delta, length)
        mid = tuple((bridge["start"][k] + bridge["end"][k]) * 0.5 for k in range(3))
        r = bridge["radius"]
        lines.append(" ".join([str(len(capsules) + 1 + i), "0", str(atom_type + 2), g(r), g(r), g(length),
                               g(q[0]), g(q[1]), g(q[2]), g(q[3]), g(mid[0]), g(mid[1]), g(mid[2]), "0", "0",
                               g(0.0), g(0.0), g(0.0), g(0.0), g(0.0), str(bridge["junction_id"])]))
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "a" if append else "w") as out:
        out.write("\n".join(lines) + "\n")
    return len(capsules) + len(bridges)


def _path_expression(path: Path, script_path: Path) -> str:
    def escape(s):
        return s.replace("\\", "\\\\").replace("'", "\\'")
    if path.parent == script_path.parent and path.name:
        return f"HERE / '{escape(path.name)}'"
    return f"Path('{escape(str(path))}')"


def write_ovito_view_script(script_path, dump_path, *, session_path=None, coloring: str = "fiber",
                            bond_radius_ratio=DEFAULT_BOND_RADIUS_RATIO, atom_type: int = 1) -> None:
    """The OVITO Python script that loads the dump, colours it and saves a
    session (write_ovito_view_script, ovito.rs:774-906, fiber-segments view)."""
    if coloring not in COLORINGS:
        raise ValueError(f"coloring must be one of {', '.join(COLORINGS)}")
    script_path, dump_path = Path(script_path), Path(dump_path)
    script_path.parent.mkdir(parents=True, exist_ok=True)
    lines = [f"# Generated by TANGLE. Run with: ovitos {script_path}", "from pathlib import Path", "import ovito",
             "from ovito.io import import_file",
             "from ovito.modifiers import AssignColorModifier, ColorCodingModifier, ExpressionSelectionModifier",
             "from ovito.vis import ParticlesVis", "", "HERE = Path(__file__).resolve().parent",
             f"trajectory_path = {_path_expression(dump_path, script_path)}",
             "pipeline = import_file(str(trajectory_path), sort_particles=True)",
             "print(f'Loaded {pipeline.num_frames} OVITO frames from {trajectory_path}')"]
    if coloring == "fiber":
        lines.append("pipeline.modifiers.append(ColorCodingModifier(property='Molecule Identifier'))")
    elif coloring == "curvature_ratio":
        lines += ["# 0 = straight; 1 = admissible bend limit; >1 = violation",
                  "pipeline.modifiers.append(ColorCodingModifier(property='curvature_ratio', start_value=0.0, "
                  "end_value=1.0))",
                  "pipeline.modifiers.append(ExpressionSelectionModifier(expression='curvature_ratio > 1.0'))",
                  "pipeline.modifiers.append(AssignColorModifier(color=(1.0, 0.0, 0.0)))"]
    else:
        lines.append("pipeline.modifiers.append(ColorCodingModifier(property='refinement_level'))")
    if bond_radius_ratio is not None:
        lines += ["# Binder bonds at junctions: one grey spherocylinder each",
                  f"pipeline.modifiers.append(ExpressionSelectionModifier(expression='ParticleType == "
                  f"{atom_type + 2}'))",
                  "pipeline.modifiers.append(AssignColorModifier(color=(0.55, 0.55, 0.55)))"]
    lines += ["pipeline.add_to_scene(name='TANGLE relaxation')", "data = pipeline.compute()",
              "data.particles.vis.shape = ParticlesVis.Shape.Spherocylinder", "data.cell.vis.enabled = True"]
    if session_path is not None:
        lines += [f"session_path = {_path_expression(Path(session_path), script_path)}",
                  "ovito.scene.save(str(session_path))", "print(f'Saved OVITO session: {session_path}')"]
    with open(script_path, "w") as out:
        out.write("\n".join(lines) + "\n")
