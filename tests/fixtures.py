"""Assemblies and settings shared by the tests and the benchmark.

The hand-built assemblies restate the fixtures of the reference's device-world
tests (crates/tangle_relax/src/device/world_tests.rs); each builder cites the
test it follows. `population_case` builds the benchmark workloads from the
seeded generator.
"""

from __future__ import annotations

import numpy as np

import tangle_amd as tg
from tangle_amd import _native


def config(**changes) -> _native.Config:
    """RelaxationConfig::default() (plugin.rs:145-175) with `changes` applied."""
    c = _native.default_config()
    for key, value in changes.items():
        setattr(c, key, value)
    return c


def packed(fibers, radius, lengths, periodic, *, rest=None, pin_fiber_ends=False, bend_radius=None,
           vertex_pinned=None, origin=(0.0, 0.0, 0.0)):
    """fibers: list of (n_i, 3) point lists. `rest` defaults to the placed points."""
    counts = [len(f) for f in fibers]
    placed = np.concatenate([np.asarray(f, dtype=np.float64) for f in fibers])
    intrinsic = placed.copy() if rest is None else np.concatenate([np.asarray(f, dtype=np.float64) for f in rest])
    radii = np.full(len(fibers), radius) if np.isscalar(radius) else np.asarray(radius)
    bend = None
    if bend_radius is not None:
        bend = np.full(len(fibers), bend_radius) if np.isscalar(bend_radius) else np.asarray(bend_radius)
    return _native.PackedAssembly(counts, placed, intrinsic, radii, bend, origin, lengths, periodic,
                                  pin_fiber_ends=pin_fiber_ends, vertex_pinned=vertex_pinned)


# world_tests.rs:8-21
def crossed(pin_fiber_ends=False):
    return packed([[[1.0, 2.0, 2.0], [3.0, 2.0, 2.0]], [[2.0, 1.0, 2.0], [2.0, 3.0, 2.0]]], 0.1, [4.0] * 3,
                  [False] * 3, pin_fiber_ends=pin_fiber_ends)


# world_tests.rs:283-300: twenty 0.05-long segments of a 0.2-diameter fiber
def short_segments():
    return packed([[[1.0 + 0.05 * i, 2.0, 2.0] for i in range(21)]], 0.1, [4.0] * 3, [False] * 3)


# world_tests.rs:303-332: a hairpin, legs 0.1 apart
def hairpin():
    pts = [[1.0 + 0.1 * i, 2.0, 2.0] for i in range(16)] + [[1.0 + 0.1 * i, 2.1, 2.0] for i in reversed(range(16))]
    return packed([pts], 0.1, [4.0] * 3, [False] * 3)


# world_tests.rs:430-458: three segments in one high-index cell of a long box
def dense_cell():
    pts = [[39.0, 0.1, 0.1], [39.01, 0.1, 0.1]]
    return packed([pts, pts, pts], 0.02, [40.0, 0.2, 0.2], [False] * 3)


# world_tests.rs:461-509 (pin_fiber_ends) and :512-554 (every vertex pinned)
def bent_triplet(pin_all=False):
    rest = [[[-0.15, 0.0, 0.0], [0.0, 0.0, 0.0], [0.15, 0.0, 0.0]]]
    placed = [[[0.8, 1.0, 1.0], [1.0, 1.15, 1.0], [1.2, 1.0, 1.0]]]
    return packed(placed, 0.01, [2.0] * 3, [False] * 3, rest=rest, pin_fiber_ends=not pin_all, bend_radius=0.25,
                  vertex_pinned=np.ones(3, dtype=np.uint8) if pin_all else None)


# world_tests.rs:934-969: minimum-image contact across a periodic x face
def periodic_pair():
    fibers = [[[x, 1.5, 2.0], [x, 2.5, 2.0]] for x in (0.05, 3.95)]
    return packed(fibers, 0.1, [4.0] * 3, [True, False, False])


# world_tests.rs:972-996: SI-scale interior crossing
def si_crossing():
    fibers = [[[0.0006, 0.001, 0.001], [0.0014, 0.001, 0.001]], [[0.001, 0.0006, 0.001], [0.001, 0.0014, 0.001]]]
    return packed(fibers, 10.0e-6, [0.002] * 3, [False] * 3)


# world_tests.rs:999-1026: a fiber several box images away
def unwrapped_images():
    fibers = [[[0.0005, 0.0006, 0.001], [0.0005, 0.0014, 0.001]], [[0.0051, 0.001, 0.001], [0.0059, 0.001, 0.001]]]
    return packed(fibers, 10.0e-6, [0.0025, 0.0025, 0.002], [True, True, False])


# world_tests.rs:1029-1064: self-contact through the periodic image
def self_image():
    pts = [[0.05, 1.0, 2.0], [0.05, 1.5, 2.0], [2.0, 1.5, 2.0], [3.95, 1.5, 2.0], [3.95, 1.0, 2.0]]
    return packed([pts], 0.1, [4.0] * 3, [True, False, False])


# world_tests.rs:1126-1157: pinned 81-vertex zigzag with a bend limit
def zigzag():
    rest = [[[0.02 * i, 0.0, 0.0] for i in range(81)]]
    placed = [[[0.2 + 0.02 * i, 1.0 + (0.0 if i % 2 == 0 else 0.01), 1.0] for i in range(81)]]
    return packed(placed, 0.005, [2.0] * 3, [False] * 3, rest=rest, pin_fiber_ends=True, bend_radius=0.3)


# world_tests.rs:1262-1287: 24 crossing fibers in a periodic-xy mat
def dense_mat():
    fibers = []
    for index in range(12):
        offset = 0.05 + 0.18 * index
        wobble = 0.03 * (index % 3)
        fibers.append([[0.2 + 0.25 * p, offset, 0.45 + wobble] for p in range(9)])
        fibers.append([[offset + 0.07, 0.2 + 0.25 * p, 0.55 - wobble] for p in range(9)])
    return packed(fibers, 0.1, [2.4, 2.4, 1.0], [True, True, False])


# Reference fixtures with the settings and expectations of their tests:
# (name, builder, config changes, iterations, expectations on the GPU status, d_min)
REFERENCE_CASES = [
    ("crossed", crossed, dict(penetration_tolerance=1e-5, correction_fraction=1.0, max_step=0.2), 20,
     lambda s: s.converged and s.total_iterations > 0 and s.max_penetration <= 1e-5, 0.2),
    ("pinned_overlap", lambda: crossed(True), dict(penetration_tolerance=1e-5, correction_fraction=1.0,
                                                      max_step=0.2), 5,
     lambda s: (not s.converged) and s.max_penetration > 1e-5, 0.2),
    ("short_segments", short_segments, dict(penetration_tolerance=1e-5, max_step=0.02), 50,
     lambda s: s.converged and s.max_penetration <= 1e-5, 0.2),
    ("hairpin", hairpin, dict(penetration_tolerance=1e-4, max_step=0.02), 4000,
     lambda s: s.max_penetration < 0.01, 0.2),
    ("dense_cell", dense_cell, dict(max_step=0.01), 5,
     lambda s: (not s.cell_list_overflow) and np.isfinite(s.max_penetration), 0.04),
    ("curvature_limit", bent_triplet, dict(penetration_tolerance=1e-5, stretch_stiffness=0.0, bend_stiffness=0.0,
                                              constraint_iterations=1, max_step=0.02), 40,
     lambda s: s.converged and s.max_curvature_ratio <= 1.0 + 1e-5, 0.02),
    ("pinned_bend", lambda: bent_triplet(pin_all=True), dict(stretch_stiffness=0.0, bend_stiffness=0.0,
                                                               constraint_iterations=1), 5,
     lambda s: (not s.converged) and s.max_curvature_ratio > 1.0 + 1e-5, 0.02),
    ("periodic_pair", periodic_pair, dict(penetration_tolerance=1e-5, max_step=0.05), 40,
     lambda s: s.converged and s.max_penetration <= 1e-5, 0.2),
    ("zigzag", zigzag, dict(penetration_tolerance=1e-5, stretch_stiffness=0.0, bend_stiffness=0.0,
                               constraint_iterations=1, curvature_cleanup_sweeps=1, max_step=0.02), 30,
     lambda s: s.converged and s.max_curvature_ratio <= 1.0 + 1e-5, 0.01),
    ("dense_mat", dense_mat, dict(penetration_tolerance=0.0, force_full_iterations=1, max_step=0.01), 40,
     lambda s: np.isfinite(s.max_penetration), 0.2),
]



# ---- benchmark workloads ---------------------------------------------------

def config1_recipe(count=40, seed=1):
    """40 straight fibers, d = 10 um, L = 300 um, 30 segments, 300 um periodic cube."""
    um = tg.um
    cell = tg.Cell([300 * um] * 3, periodic="xyz")
    fiber = tg.Material("fiber", diameter=10 * um)
    pop = tg.FiberPopulation(material=fiber, count=count, length=300 * um, segments_per_fiber=30,
                             curvature_amplitude=0.0, seed=seed)
    recipe = tg.Recipe(cell)
    recipe.insert(tg.generate_fiber_population(cell, pop))
    return recipe


# The bimodal felt specimen: 1,763 fibers of 7 um and 237 of 19 um (equal
# volume), 1 mm long in a 1 mm periodic cube, segment length ~2 d, planar
# bias about z. Schladitz beta = 1/0.3 in the reference's fit convention
# (beta > 1 lays fibers in the plane). Bend limits and seeds follow the
# reference's felt example (examples/felt_20ply_control_vs_needled/src/recipe.rs:397-420).
BIMODAL_FINE = dict(diameter=7e-6, count=1763, segments=71, bend=200e-6, seed=20_260_921)
BIMODAL_COARSE = dict(diameter=19e-6, count=237, segments=26, bend=60e-6, seed=20_260_922)


def bimodal_recipe(scale=1.0, box=1e-3, length=1e-3, beta=1.0 / 0.3, seed_offset=0):
    """scale < 1 keeps the volume fraction and shrinks the box and fiber counts;
    `seed_offset` draws another specimen of the same distribution."""
    side = box * scale ** (1.0 / 3.0) if scale != 1.0 else box
    cell = tg.Cell([side] * 3, periodic="xyz")
    recipe = tg.Recipe(cell)
    for spec in (BIMODAL_FINE, BIMODAL_COARSE):
        material = tg.Material("fine" if spec is BIMODAL_FINE else "coarse", spec["diameter"], spec["bend"])
        count = max(1, int(round(spec["count"] * scale)))
        flen = min(length, side)
        segments = max(2, int(round(spec["segments"] * flen / length)))
        pop = tg.FiberPopulation(material=material, count=count, length=flen, segments_per_fiber=segments,
                                 seed=spec["seed"] + seed_offset, orientation=tg.SchladitzOrientation(beta=beta, axis="z"))
        recipe.insert(tg.generate_fiber_population(cell, pop))
    return recipe


def hard_gate_settings(d_min, **changes):
    """Hard admissibility: penetration <= 1e-3 d_min, curvature ratio <= 1 + 1e-5."""
    return tg.RelaxationSettings(penetration_tolerance=1e-3 * d_min, curvature_ratio_tolerance=1e-5, **changes)


HARD_PENETRATION = 1e-3 * 7e-6   # 1e-3 d_min
HARD_CURVATURE_RATIO = 1.0 + 1e-5


def staged_hard_gate_recipe(recipe, warm_iterations=1280, cleanup_iterations=40000, preset="contact_cleanup"):
    """The plain path's plan to the hard gates, in the reference's recipe
    vocabulary: `warm_iterations` of default relaxation, then a final
    SolvePolicy at the hard limits under a cleanup preset (as the reference's
    tutorials finish their felts, crates/tangle_python/python/tutorials/generate.py:745-753)."""
    if warm_iterations:
        recipe.relax_for(warm_iterations)
    policy = tg.SolvePolicy("final", target_penetration=HARD_PENETRATION, target_curvature_ratio=HARD_CURVATURE_RATIO,
                            max_iterations=cleanup_iterations, max_extra_iterations=0)
    recipe.solve(policy, tg.RelaxationOverrides.preset(preset) if preset else None)
    return recipe


# Batched specimens: independent cubes of straight-ish 10 um fibers 0.5 mm
# long (25 segments of 2 d), isotropic, each cube sized for 10 % fill, with
# fiber counts log-uniform in [n_min, n_max] (seeded).
def batched_specimen_recipes(count=256, n_min=50, n_max=5000, seed=20_260_930, fill=0.10):
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count)).round().astype(int)
    d, length = 10e-6, 0.5e-3
    recipes = []
    for i, n in enumerate(counts):
        side = (n * np.pi * (d / 2) ** 2 * length / fill) ** (1.0 / 3.0)
        cell = tg.Cell([side] * 3, periodic="xyz")
        pop = tg.FiberPopulation(material=tg.Material("fiber", d), count=int(n), length=length,
                                 segments_per_fiber=25, seed=seed + 1 + i)
        recipe = tg.Recipe(cell)
        recipe.insert(tg.generate_fiber_population(cell, pop))
        recipes.append(recipe)
    return recipes


# The bimodal mix at scale: 20,000 fibers (17,630 x 7 um + 2,370 x 19 um)
# 1 mm long in a 3 x 3 x 1 mm slab, periodic in x and y, walled in z.
def slab_recipe(fibers=20_000, beta=1.0 / 0.3, seed_offset=0):
    cell = tg.Cell([3e-3, 3e-3, 1e-3], periodic="xy")
    recipe = tg.Recipe(cell)
    total = BIMODAL_FINE["count"] + BIMODAL_COARSE["count"]
    for spec in (BIMODAL_FINE, BIMODAL_COARSE):
        material = tg.Material("fine" if spec is BIMODAL_FINE else "coarse", spec["diameter"], spec["bend"])
        count = int(round(spec["count"] * fibers / total))
        pop = tg.FiberPopulation(material=material, count=count, length=1e-3, segments_per_fiber=spec["segments"],
                                 seed=spec["seed"] + 1000 + seed_offset,
                                 orientation=tg.SchladitzOrientation(beta=beta, axis="z"))
        recipe.insert(tg.generate_fiber_population(cell, pop))
    return recipe


# The layered, needled felt: plies of the bimodal mix (2,201 x 7 um + 299 x
# 19 um each, 1 mm long, segment ~2 d, bend limits as above) on a 2 x 2 mm
# periodic footprint, bounded along z, each ply generated in its own band
# `gap` thick. Each ply is inserted as a formation step and relaxed; then
# every ply has 5 % of its 19 um fibers pulled 30 um down through the stack,
# held until the needles reach their targets, and released; a final cleanup
# (contact_cleanup preset, fixed budget, soft gates) and the post-recipe
# relaxation to the run's tolerances follow. With `place`, each new ply is
# also held `gap` above the last while it relaxes (place_layer_above). The
# reference's analogue is its 20-ply felt example
# (examples/felt_20ply_control_vs_needled/src/main.rs:42-58, recipe.rs:9-66),
# whose plies hold about 27 fibers each; a layer target pulls every fiber's
# centre to its ply's plane, which for plies of 2,500 fibers (a 42 um solid
# sheet) crushes the ply into deep crossings that do not come apart, so the
# benchmark deposits plies in their bands instead. Compaction is left out.
def layered_felt_recipe(plies=20, fine=2201, coarse=299, footprint=2e-3, gap=300e-6, relax_per_ply=64,
                        needle_fraction=0.05, needle_depth=30e-6, settle_iterations=64, final_iterations=512,
                        seed=20_260_921, place=False, needle=True):
    import math
    height = plies * gap
    cell = tg.Cell([footprint, footprint, height], periodic="xy")
    population = None
    for name, d, count, bend, offset in (("fine", 7e-6, fine, 200e-6, 0), ("coarse", 19e-6, coarse, 60e-6, 1)):
        material = tg.Material(name, d, bend)
        pop = tg.FiberPopulation(material=material, count=count * plies, length=1e-3,
                                 segments_per_fiber=max(2, round(1e-3 / (2 * d))), seed=seed + offset,
                                 curvature_amplitude=(0.0, 0.4 * d),
                                 orientation=tg.PlanarOrientation(normal="z", max_tilt=math.radians(3.0)),
                                 position=tg.LayeredPosition(plies, jitter_fraction=1.0))
        col = tg.generate_fiber_population(cell, pop, name=name)
        population = col if population is None else population + col
    recipe = tg.Recipe(cell)
    for ply in range(plies):
        recipe.insert(population.select_layer(ply), name=f"ply {ply}")
        if ply == 0 or not place:
            recipe.relax_for(relax_per_ply)
        else:
            with recipe.place_layer_above(ply, gap=gap, stiffness=1.0, max_translation=2e-6):
                recipe.relax_for(relax_per_ply)
    for ply in range(plies if needle else 0):
        with recipe.needle_layer(ply, footprint=tg.RandomFiberFraction(needle_fraction, seed=seed + 7 + ply),
                                 depth=needle_depth, min_fiber_diameter=19e-6, stiffness=1.0, max_translation=2e-6):
            recipe.settle_targets(tolerance=0.05 * needle_depth, max_iterations=settle_iterations)
    # final contact and curvature cleanup toward the hard gates, for a fixed
    # budget: its limits are soft, so an unfinished cleanup is reported, not raised
    policy = tg.SolvePolicy("final cleanup", target_penetration=HARD_PENETRATION,
                            target_curvature_ratio=HARD_CURVATURE_RATIO, max_iterations=final_iterations,
                            max_extra_iterations=0, hard_penetration=False, hard_curvature=False,
                            on_budget_exhausted="continue_if_hard_ok")
    recipe.solve(policy, tg.RelaxationOverrides.preset("contact_cleanup"))
    settings = tg.RelaxationSettings(max_iterations=plies * (relax_per_ply + settle_iterations) + final_iterations,
                                     iterations_per_batch=64)
    return recipe, settings
