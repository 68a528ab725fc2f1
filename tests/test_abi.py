"""The C ABI on the CPU: the built library loads (no device call is made),
exports every function include/tangle_hip.h declares, and the Python binding
declares a signature for each; the host facade validates its arguments as
the reference's Python layer does (settings.rs:500-567)."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

import tangle_amd as tg
from tangle_amd import _native

HEADER = Path(__file__).resolve().parent.parent / "include" / "tangle_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(tgh_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("tgh_world_create", "tgh_world_run_batch", "tgh_world_download_positions",
                     "tgh_world_info_get", "tgh_world_destroy", "tgh_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = _native.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    missing = [n for n in declared_functions() if n not in _native.SIGNATURES]
    assert not missing, missing


def test_config_default_matches_reference_defaults():
    c = _native.default_config()  # no device call
    # RelaxationConfig::default() (plugin.rs:145-175)
    assert c.correction_fraction == pytest.approx(0.8)
    assert c.stretch_stiffness == pytest.approx(0.35)
    assert c.bend_stiffness == pytest.approx(0.03)
    assert c.constraint_iterations == 2
    assert c.curvature_cleanup_sweeps == 4
    assert c.neighbor_capacity == 48


def test_mlp_parameter_count_is_the_layer_sizes():
    assert _native.library().tgh_mlp_parameter_count() == 64 * 64 + 64 + 64 * 64 + 64 + 16 * 64 + 16


@pytest.mark.parametrize("changes", [dict(backend="cpu"), dict(correction_fraction=0.0),
                                     dict(stretch_stiffness=1.5), dict(curvature_limit_safety_margin=0.2),
                                     dict(max_iterations=0), dict(cell_size_scale=0.5),
                                     dict(motion_model="rigid_translation")])
def test_settings_reject_invalid_values(changes):
    with pytest.raises(ValueError):
        tg.RelaxationSettings(**changes)


def test_recipe_rejects_empty_runs():
    recipe = tg.Recipe(tg.Cell([1e-3] * 3, periodic="xyz"))
    with pytest.raises(ValueError):
        recipe.run()
    with pytest.raises(ValueError):
        recipe.relax_for(0)


def test_presets_follow_the_reference():
    # RelaxationOverrides presets (tangle_python/src/recipe.rs:1123-1267)
    p = tg.RelaxationOverrides.preset("contact_cleanup")
    assert p.correction_fraction == pytest.approx(0.8)
    assert p.curvature_cleanup_sweeps == 16
    with pytest.raises(ValueError):
        tg.RelaxationOverrides.preset("no_such_preset")


@pytest.mark.parametrize("case,message", [
    ("empty", "no fibers"), ("single_vertex", "fewer than two vertices"), ("nan", "finite"),
    ("zero_radius", "radius must be positive"), ("zero_cell", "cell lengths"),
])
def test_world_creation_rejects_invalid_assemblies(case, message):
    """Invalid assemblies are rejected by the C ABI before any device call
    (so this runs without a GPU)."""
    import numpy as np

    import fixtures as fx
    from tangle_amd import _native

    two = [[0.5, 0.5, 0.5], [0.6, 0.5, 0.5]]
    if case == "empty":
        pk = _native.PackedAssembly([], np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), None, [0, 0, 0], [1, 1, 1],
                                    [False] * 3)
    elif case == "single_vertex":
        pk = fx.packed([[[0.5, 0.5, 0.5]]], 0.1, [1.0] * 3, [False] * 3)
    elif case == "nan":
        pk = fx.packed([[[0.5, 0.5, float("nan")], [0.6, 0.5, 0.5]]], 0.1, [1.0] * 3, [False] * 3)
    elif case == "zero_radius":
        pk = fx.packed([two], 0.0, [1.0] * 3, [False] * 3)
    else:
        pk = fx.packed([two], 0.1, [1.0, 0.0, 1.0], [False] * 3)
    with pytest.raises(ValueError, match=message):
        _native.World(pk, fx.config())
