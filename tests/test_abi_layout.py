"""The C ABI's struct layouts, checked three ways:
- the header compiled by gcc (sizeof, offsetof of every field) against the
  ctypes structures the Python binding passes;
- the Rust #[repr(C)] structs that INTEGRATION.md gives a maintainer to
  copy: same fields in the same order, and the offsets C layout rules give
  their Rust types equal to gcc's;
- tgh_world_create_packed rejects layouts the backend does not relax before
  any device call.
"""

from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fixtures as fx
from tangle_amd import _native

ROOT = Path(__file__).resolve().parent.parent

STRUCTS = {
    "tgh_assembly": _native.Assembly,
    "tgh_packed": _native.PackedStruct,
    "tgh_config": _native.Config,
    "tgh_batch_status": _native.BatchStatus,
    "tgh_world_info": _native.WorldInfo,
    "tgh_tuning": _native.Tuning,
    "tgh_device_info": _native.DeviceInfo,
    "tgh_checkpoint": _native.CheckpointStruct,
    "tgh_compaction_metrics": _native.CompactionMetricsStruct,
    "tgh_contact_candidate": _native.ContactCandidateStruct,
    "tgh_population_spec": _native.PopulationSpec,
}
RUST = {"TghPacked": "tgh_packed", "TghAssembly": "tgh_assembly", "TghConfig": "tgh_config",
        "TghBatchStatus": "tgh_batch_status", "TghCompactionMetrics": "tgh_compaction_metrics"}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tangle_hip.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines))
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    out = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        out[(cname, field)] = int(value)
    return out


@pytest.mark.parametrize("cname", list(STRUCTS))
def test_ctypes_matches_the_header(c_layout, cname):
    cls = STRUCTS[cname]
    assert c_layout[(cname, "sizeof")] == __import__("ctypes").sizeof(cls), cname
    for field, _ in cls._fields_:
        assert c_layout[(cname, field)] == getattr(cls, field).offset, (cname, field)


RUST_SCALARS = {"u8": 1, "u16": 2, "u32": 4, "i32": 4, "f32": 4, "u64": 8, "f64": 8, "usize": 8}


def rust_size_align(ty: str):
    ty = ty.strip()
    if ty.startswith("*const") or ty.startswith("*mut"):
        return 8, 8
    m = re.fullmatch(r"\[(.+);\s*(\d+)\]", ty)
    if m:
        size, align = rust_size_align(m.group(1))
        return size * int(m.group(2)), align
    return RUST_SCALARS[ty], RUST_SCALARS[ty]


def rust_structs():
    text = (ROOT / "INTEGRATION.md").read_text()
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\][^\n{]*pub struct (\w+)\s*\{(.*?)\n?\}", text, re.S):
        fields = re.findall(r"pub (\w+):\s*([^,]+(?:\[[^\]]*\])?[^,]*),", m.group(2) + ",")
        out[m.group(1)] = [(n, t.strip()) for n, t in fields]
    return out


@pytest.mark.parametrize("rname", list(RUST))
def test_integration_rust_structs_match_the_header(c_layout, rname):
    structs = rust_structs()
    assert rname in structs, f"INTEGRATION.md has no #[repr(C)] struct {rname}"
    cname = RUST[rname]
    fields = structs[rname]
    assert [n for n, _ in fields] == [n for n, _ in STRUCTS[cname]._fields_], rname
    offset, max_align = 0, 1
    for name, ty in fields:
        size, align = rust_size_align(ty)
        offset = (offset + align - 1) // align * align
        assert offset == c_layout[(cname, name)], (rname, name, offset)
        offset += size
        max_align = max(max_align, align)
    assert (offset + max_align - 1) // max_align * max_align == c_layout[(cname, "sizeof")], rname


def _layout(**changes):
    packed = fx.bimodal_recipe(scale=0.002)._packed(False)
    base = _native.PackedLayout.from_assembly(packed)
    args = dict(positions=base.positions, intrinsic_positions=base.intrinsic_positions,
                segment_vertices=base.segment_vertices, segment_fibers=base.segment_fibers,
                segment_radii=base.segment_radii, segment_rest_lengths=base.segment_rest_lengths,
                fiber_vertex_spans=base.fiber_vertex_spans, vertex_max_curvature=base.vertex_max_curvature,
                vertex_pinned=base.vertex_pinned, cell_lower=list(base.struct.cell_lower),
                cell_upper=list(base.struct.cell_upper), cell_periodic=list(base.struct.cell_periodic))
    extra = {k: changes.pop(k) for k in ("segment_active", "adaptive", "has_ovals") if k in changes}
    for k, v in changes.items():
        args[k] = v(args[k].copy()) if callable(v) else v
    return _native.PackedLayout(**args, **extra)


def _spans_gap(spans):
    spans = spans.reshape(-1, 2)
    spans[1, 0] += 1
    return spans


def _swap_segment(sv):
    sv = sv.reshape(-1, 2)
    sv[3] = sv[3][::-1]
    return sv


def _radius_step(r):
    r[1] *= 1.5
    return r


@pytest.mark.parametrize("changes,message", [
    (dict(adaptive=True), "adaptive"),
    (dict(has_ovals=True), "oval"),
    (dict(fiber_vertex_spans=_spans_gap), "span"),
    (dict(segment_vertices=_swap_segment), "uniform layout"),
    (dict(segment_radii=_radius_step), "radius"),
    (dict(segment_active="inactive"), "inactive"),
    (dict(cell_upper=[0.0, 1.0, 1.0]), "cell bounds"),
])
def test_create_packed_rejects_other_layouts(changes, message):
    if changes.get("segment_active") == "inactive":
        n = _layout().segment_fibers.size
        active = np.ones(n, dtype=np.uint32)
        active[5] = 0
        changes = dict(segment_active=active)
    layout = _layout(**changes)
    with pytest.raises(ValueError, match=message):
        _native.World.from_packed(layout, fx.config())


def test_packed_layout_follows_from_assembly():
    # a two-fiber assembly whose packing is known by hand
    packed = fx.packed([[[0.0, 0.0, 0.0], [3.0, 4.0, 0.0], [3.0, 4.0, 1.0]], [[1.0, 1.0, 1.0], [1.0, 1.0, 2.0]]],
                       [0.5, 0.25], [10.0] * 3, [True, True, False], bend_radius=[2.0, 4.0], pin_fiber_ends=True)
    layout = _native.PackedLayout.from_assembly(packed)
    assert layout.segment_vertices.reshape(-1, 2).tolist() == [[0, 1], [1, 2], [3, 4]]
    assert layout.segment_fibers.tolist() == [0, 0, 1]
    assert layout.segment_radii.tolist() == [0.5, 0.5, 0.25]
    assert layout.segment_rest_lengths.tolist() == [5.0, 1.0, 1.0]
    assert layout.fiber_vertex_spans.reshape(-1, 2).tolist() == [[0, 3], [3, 2]]
    # the bend limit sits on internal vertices only (packed.rs:260-270)
    assert layout.vertex_max_curvature.tolist() == [0.0, 0.5, 0.0, 0.0, 0.0]
    assert layout.vertex_pinned.tolist() == [1, 0, 1, 1, 1]
    assert list(layout.struct.cell_periodic) == [1, 1, 0]
