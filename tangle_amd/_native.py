"""ctypes binding of libtangle_hip.so (the C ABI in include/tangle_hip.h).

The library is required: importing a function here that needs it raises
RuntimeError when it has not been built (``python -m tangle_amd.build``).
There is no fallback implementation.
"""

from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "_lib" / "libtangle_hip.so"

TGH_OK = 0
TGH_ERR_INVALID = -1
TGH_ERR_DEVICE = -2
TGH_ERR_SIZE = -3
TGH_MLP_FORWARD_STATIC_MAX_ROWS = 65536  # include/tangle_hip.h

ORIENT_ISOTROPIC = 0
ORIENT_PLANAR = 1
ORIENT_ALIGNED = 2
ORIENT_SCHLADITZ = 3
ORIENT_LAYERED_BIAXIAL = 4
POSITION_UNIFORM = 0
POSITION_LAYERED = 1
POSITION_GRADIENT = 2

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)


class Assembly(C.Structure):
    _fields_ = [
        ("fiber_count", C.c_uint32),
        ("vertex_count", C.c_uint32),
        ("fiber_vertex_counts", _u32p),
        ("placed", _f64p),
        ("intrinsic", _f64p),
        ("fiber_radius", _f64p),
        ("fiber_min_bend_radius", _f64p),
        ("vertex_pinned", _u8p),
        ("cell_origin", C.c_double * 3),
        ("cell_lengths", C.c_double * 3),
        ("cell_periodic", C.c_uint32 * 3),
        ("pin_fiber_ends", C.c_uint32),
        ("fiber_formation_step", C.POINTER(C.c_uint32)),
        ("fiber_formation_layer", C.POINTER(C.c_uint32)),
    ]


class PackedStruct(C.Structure):
    """tgh_packed: PackedAssembly's f32 arrays (packed.rs:47-141)."""
    _fields_ = [
        ("fiber_count", C.c_uint32),
        ("vertex_count", C.c_uint32),
        ("segment_count", C.c_uint32),
        ("positions", _f32p),
        ("intrinsic_positions", _f32p),
        ("segment_vertices", _u32p),
        ("segment_fibers", _u32p),
        ("segment_radii", _f32p),
        ("segment_rest_lengths", _f32p),
        ("segment_active", _u32p),
        ("fiber_vertex_spans", _u32p),
        ("fiber_formation_layers", _u32p),
        ("fiber_formation_steps", _u32p),
        ("vertex_max_curvature", _f32p),
        ("vertex_pinned", _u32p),
        ("cell_lower", C.c_float * 3),
        ("cell_upper", C.c_float * 3),
        ("cell_periodic", C.c_uint32 * 3),
        ("adaptive", C.c_uint32),
        ("has_ovals", C.c_uint32),
    ]


class Config(C.Structure):
    _fields_ = [
        ("penetration_tolerance", C.c_float),
        ("correction_fraction", C.c_float),
        ("stretch_stiffness", C.c_float),
        ("bend_stiffness", C.c_float),
        ("curvature_limit_stiffness", C.c_float),
        ("curvature_limit_safety_margin", C.c_float),
        ("curvature_ratio_tolerance", C.c_float),
        ("max_step", C.c_float),
        ("cell_size_scale", C.c_float),
        ("neighbor_skin_scale", C.c_float),
        ("contact_aggregation", C.c_uint32),
        ("constraint_iterations", C.c_uint32),
        ("curvature_cleanup_sweeps", C.c_uint32),
        ("force_full_iterations", C.c_uint32),
        ("neighbor_capacity", C.c_uint32),
        ("maximum_neighbor_list_bytes", C.c_uint32),
    ]


class BatchStatus(C.Structure):
    _fields_ = [
        ("total_iterations", C.c_uint64),
        ("batch_iterations", C.c_uint64),
        ("max_penetration", C.c_float),
        ("max_displacement", C.c_float),
        ("max_curvature_ratio", C.c_float),
        ("converged", C.c_uint32),
        ("cell_list_overflow", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class WorldInfo(C.Structure):
    _fields_ = [
        ("fiber_count", C.c_uint32),
        ("segment_count", C.c_uint32),
        ("vertex_count", C.c_uint32),
        ("cells", C.c_uint32 * 3),
        ("cell_count", C.c_uint32),
        ("cell_size", C.c_float),
        ("neighbor_skin", C.c_float),
        ("neighbor_block", C.c_uint32),
        ("neighbor_list_rebuilds", C.c_uint32),
        ("device_bytes", C.c_uint64),
        ("specimen_count", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("curvature_lanes", C.c_uint32),
        ("contact_lanes", C.c_uint32),
        ("reserved", C.c_uint32 * 6),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256),
        ("arch", C.c_char * 64),
        ("compute_units", C.c_uint32),
        ("reserved", C.c_uint32 * 7),
    ]


class CheckpointStruct(C.Structure):
    _fields_ = [
        ("total_iterations", C.c_uint64),
        ("fiber_count", C.c_uint32),
        ("vertex_count", C.c_uint32),
        ("positions", C.POINTER(C.c_float)),
        ("fiber_active", C.POINTER(C.c_uint32)),
        ("wall_reactions", C.POINTER(C.c_float)),
        ("layer_target_count", C.c_uint32),
        ("layer_axis", C.c_uint32),
        ("layer_first", C.c_uint32),
        ("layer_last", C.c_uint32),
        ("layer_stiffness", C.c_float),
        ("layer_max_translation", C.c_float),
        ("layer_targets", C.POINTER(C.c_float)),
        ("vertex_target_count", C.c_uint32),
        ("vertex_axis", C.c_uint32),
        ("vertex_stiffness", C.c_float),
        ("vertex_max_translation", C.c_float),
        ("vertex_indices", C.POINTER(C.c_uint32)),
        ("vertex_targets", C.POINTER(C.c_float)),
        ("vertex_commands", C.POINTER(C.c_float)),
        ("cell_lower", C.c_float * 3),
        ("cell_upper", C.c_float * 3),
    ]


class CompactionMetricsStruct(C.Structure):
    _fields_ = [
        ("pressure", C.c_float * 3),
        ("contact_energy", C.c_float),
        ("stretch_energy", C.c_float),
        ("bending_energy", C.c_float),
        ("total_energy", C.c_float),
        ("face_pressure", (C.c_float * 2) * 3),
    ]


class ContactCandidateStruct(C.Structure):
    _fields_ = [
        ("first_segment", C.c_uint32),
        ("second_segment", C.c_uint32),
        ("first_coordinate", C.c_float),
        ("second_coordinate", C.c_float),
        ("surface_gap", C.c_float),
        ("crossing_angle", C.c_float),
    ]


# CompactionKinematics (crates/tangle_relax/src/compaction.rs), by the names
# the Python API uses (crates/tangle_python/src/compaction.rs:15-22)
COMPACT_RIGID_FIBER_CENTERS = 0
COMPACT_MOVING_WALLS = 1
COMPACT_AFFINE_VERTICES = 2
KINEMATICS = {"rigid_fiber_centers": COMPACT_RIGID_FIBER_CENTERS, "moving_walls": COMPACT_MOVING_WALLS,
              "affine_vertices": COMPACT_AFFINE_VERTICES}


def kinematics_code(kinematics) -> int:
    if isinstance(kinematics, str):
        if kinematics not in KINEMATICS:
            raise ValueError(f"kinematics must be one of {', '.join(KINEMATICS)}")
        return KINEMATICS[kinematics]
    code = int(kinematics)
    if code not in KINEMATICS.values():
        raise ValueError("unknown compaction kinematics")
    return code


@dataclass(frozen=True)
class CompactionEnergyModel:
    """Penalty stiffnesses of the compaction metrics (CompactionEnergyModel,
    crates/tangle_relax/src/compaction.rs:21-38)."""

    contact_stiffness: float = 1.0
    stretch_stiffness: float = 1.0
    bending_stiffness: float = 1.0


@dataclass(frozen=True)
class CompactionMetrics:
    """Directional pressures and penalty energies of the resident state
    (CompactionMetrics, compaction.rs:42-53), plus the lower/upper face
    pressures per axis (wall_face_pressures, world.rs:2113-2127)."""

    pressure: tuple
    contact_energy: float
    stretch_energy: float
    bending_energy: float
    total_energy: float
    face_pressures: tuple  # ((lower x, upper x), (lower y, upper y), (lower z, upper z))


@dataclass(frozen=True)
class SegmentContactCandidate:
    """One captured inter-fiber segment pair (SegmentContactCandidate,
    world.rs:44-57)."""

    first_segment: int
    second_segment: int
    first_coordinate: float
    second_coordinate: float
    surface_gap: float
    crossing_angle: float


@dataclass(frozen=True)
class ContactCapture:
    """Captured pairs, sorted, and whether the capacity was exceeded
    (ContactCapture, world.rs:61-66)."""

    candidates: list
    overflow: bool


@dataclass
class DeviceCheckpoint:
    """The persistent state of a one-specimen world (DeviceWorldCheckpoint,
    crates/tangle_relax/src/device/world.rs:281-304): positions, fiber
    activity, wall reactions, the lifetime iteration count and the held
    formation targets (None when not held)."""

    total_iterations: int
    positions: np.ndarray        # (3V,) float32
    fiber_active: np.ndarray     # (F,) uint32
    wall_reactions: np.ndarray   # (3V,) float32
    layer_target: dict | None = None   # axis, first, last, stiffness, max_translation, targets
    vertex_target: dict | None = None  # axis, stiffness, max_translation, indices, targets, commands
    cell_lower: tuple | None = None    # the world's cell when compacted since creation
    cell_upper: tuple | None = None

    def struct(self) -> CheckpointStruct:
        """A tgh_checkpoint view of these arrays (keep self alive while used)."""
        c = CheckpointStruct()
        c.total_iterations = int(self.total_iterations)
        c.fiber_count = int(self.fiber_active.size)
        c.vertex_count = int(self.positions.size // 3)
        self.positions = np.ascontiguousarray(self.positions, np.float32)
        self.fiber_active = np.ascontiguousarray(self.fiber_active, np.uint32)
        self.wall_reactions = np.ascontiguousarray(self.wall_reactions, np.float32)
        c.positions = ptr(self.positions, C.c_float)
        c.fiber_active = ptr(self.fiber_active, C.c_uint32)
        c.wall_reactions = ptr(self.wall_reactions, C.c_float)
        for k in range(3):  # NaN: the cell of the assembly the world is made from
            c.cell_lower[k] = float("nan") if self.cell_lower is None else float(self.cell_lower[k])
            c.cell_upper[k] = float("nan") if self.cell_upper is None else float(self.cell_upper[k])
        if self.layer_target is not None:
            lt = self.layer_target
            lt["targets"] = np.ascontiguousarray(lt["targets"], np.float32)
            c.layer_target_count = lt["targets"].size
            c.layer_axis, c.layer_first, c.layer_last = lt["axis"], lt["first"], lt["last"]
            c.layer_stiffness, c.layer_max_translation = lt["stiffness"], lt["max_translation"]
            c.layer_targets = ptr(lt["targets"], C.c_float)
        if self.vertex_target is not None:
            vt = self.vertex_target
            for key, dt in (("indices", np.uint32), ("targets", np.float32), ("commands", np.float32)):
                vt[key] = np.ascontiguousarray(vt[key], dt)
            c.vertex_target_count = vt["indices"].size
            c.vertex_axis = vt["axis"]
            c.vertex_stiffness, c.vertex_max_translation = vt["stiffness"], vt["max_translation"]
            c.vertex_indices = ptr(vt["indices"], C.c_uint32)
            c.vertex_targets = ptr(vt["targets"], C.c_float)
            c.vertex_commands = ptr(vt["commands"], C.c_float)
        return c


class PopulationSpec(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("segments_per_fiber", C.c_uint32),
        ("seed", C.c_uint64),
        ("length_min", C.c_double),
        ("length_max", C.c_double),
        ("radius_min", C.c_double),
        ("radius_max", C.c_double),
        ("amplitude_min", C.c_double),
        ("amplitude_max", C.c_double),
        ("orientation", C.c_uint32),
        ("orientation_vector", C.c_double * 3),
        ("orientation_angle", C.c_double),
        ("orientation_beta", C.c_double),
        ("min_bend_radius", C.c_double),
        ("max_attempts_per_fiber", C.c_uint32),
        ("cell_origin", C.c_double * 3),
        ("cell_lengths", C.c_double * 3),
        ("cell_periodic", C.c_uint32 * 3),
        ("position", C.c_uint32),
        ("position_axis", C.c_uint32),
        ("position_layers", C.c_uint32),
        ("position_toward_high", C.c_uint32),
        ("position_jitter_fraction", C.c_double),
        ("position_exponent", C.c_double),
        ("orientation_primary_fraction", C.c_double),
        ("orientation_cross_fraction", C.c_double),
        ("orientation_in_plane_deviation", C.c_double),
        ("orientation_layer_seed", C.c_uint64),
    ]


# Every symbol the header declares, with (restype, argtypes).
SIGNATURES = {
    "tgh_config_default": (None, [C.POINTER(Config)]),
    "tgh_world_create": (C.c_int, [C.POINTER(Assembly), C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "tgh_world_create_packed": (C.c_int, [C.POINTER(PackedStruct), C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "tgh_world_create_batch": (C.c_int, [C.POINTER(Assembly), C.c_uint32, C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "tgh_world_specimen_status": (C.c_int, [C.c_void_p, C.POINTER(BatchStatus), C.c_size_t]),
    "tgh_world_run_batch": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_uint32, C.POINTER(BatchStatus)]),
    "tgh_world_download_positions": (C.c_int, [C.c_void_p, _f32p, C.c_size_t]),
    "tgh_world_upload_positions": (C.c_int, [C.c_void_p, _f32p, C.c_size_t]),
    "tgh_world_info_get": (C.c_int, [C.c_void_p, C.POINTER(WorldInfo)]),
    "tgh_world_rebuild_lists": (C.c_int, [C.c_void_p]),
    "tgh_world_cell_list": (C.c_int, [C.c_void_p, _u32p, _u32p, _u32p, C.c_size_t, C.c_size_t]),
    "tgh_world_neighbor_lists": (C.c_int, [C.c_void_p, _u32p, _u32p, C.c_size_t]),
    "tgh_world_set_timing": (C.c_int, [C.c_void_p, C.c_uint32]),
    "tgh_world_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "tgh_world_activate_formation_step": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64),
                                                    C.POINTER(C.c_uint64)]),
    "tgh_world_set_layer_targets": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, C.c_size_t, C.c_uint32, C.c_uint32,
                                              C.c_float, C.c_float]),
    "tgh_world_clear_layer_targets": (C.c_int, [C.c_void_p]),
    "tgh_world_set_vertex_targets": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_size_t, C.c_float,
                                               C.c_float, C.c_float]),
    "tgh_world_clear_vertex_targets": (C.c_int, [C.c_void_p]),
    "tgh_world_formation_target_error": (C.c_int, [C.c_void_p, _f32p, _f32p, C.POINTER(C.c_uint32),
                                                   C.POINTER(C.c_uint32)]),
    "tgh_world_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "tgh_device_info_get": (C.c_int, [C.POINTER(DeviceInfo)]),
    "tgh_world_checkpoint_sizes": (C.c_int, [C.c_void_p, C.POINTER(CheckpointStruct)]),
    "tgh_world_cell_bounds": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "tgh_world_compact_cell": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32]),
    "tgh_world_compaction_metrics": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                                               C.POINTER(CompactionMetricsStruct)]),
    "tgh_world_capture_contacts": (C.c_int, [C.c_void_p, C.c_float, C.c_size_t, C.POINTER(ContactCandidateStruct),
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "tgh_world_set_fiber_maximum_curvatures": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_size_t]),
    "tgh_world_checkpoint": (C.c_int, [C.c_void_p, C.POINTER(CheckpointStruct)]),
    "tgh_world_restore": (C.c_int, [C.POINTER(Assembly), C.POINTER(Config), C.POINTER(CheckpointStruct),
                                    C.POINTER(C.c_void_p)]),
    "tgh_world_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_size_t, C.c_uint32]),
    "tgh_world_destroy": (None, [C.c_void_p]),
    "tgh_last_error": (C.c_char_p, []),
    "tgh_generate_population": (C.c_int, [C.POINTER(PopulationSpec), _f64p, _f64p, _f64p, C.POINTER(C.c_uint32),
                                          C.c_size_t]),
    "tgh_mlp_parameter_count": (C.c_size_t, []),
    "tgh_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "tgh_mlp_destroy": (None, [C.c_void_p]),
    "tgh_mlp_get_parameters": (C.c_int, [C.c_void_p, _f32p, C.c_size_t]),
    "tgh_mlp_set_parameters": (C.c_int, [C.c_void_p, _f32p, C.c_size_t]),
    "tgh_mlp_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), C.c_size_t, _f32p]),
    "tgh_mlp_forward_static": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), C.c_size_t, _f32p]),
    "tgh_mlp_gradients": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), _f32p, C.c_size_t, _f32p, _f32p]),
    "tgh_mlp_record_begin": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tgh_mlp_record_end": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tgh_mlp_dataset_rows": (C.c_uint64, [C.c_void_p]),
    "tgh_mlp_dataset_clear": (C.c_int, [C.c_void_p]),
    "tgh_mlp_train": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _f32p]),
    "tgh_world_features": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), C.c_size_t]),
    "tgh_world_mark_history": (C.c_int, [C.c_void_p]),
    "tgh_mlp_time_kernels": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "tgh_world_propose": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float]),
    "tgh_world_revert_proposal": (C.c_int, [C.c_void_p]),
    "tgh_world_revert_proposal_specimens": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_size_t]),
    "tgh_mfma_probe": (C.c_int, [C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), _f32p]),
}

_lock = threading.Lock()
_lib = None


def library() -> C.CDLL:
    """Loads libtangle_hip.so once; raises RuntimeError when it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"{LIBRARY_PATH} is not built; run `python -m tangle_amd.build` "
                    "(the HIP backend has no fallback)"
                )
            lib = C.CDLL(str(LIBRARY_PATH))
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
    return _lib


class NativeError(Exception):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def device_info() -> dict:
    """The current HIP device: name, architecture and compute units."""
    info = DeviceInfo()
    check(library().tgh_device_info_get(C.byref(info)))
    return {"name": info.name.decode(errors="replace"), "arch": info.arch.decode(errors="replace"),
            "compute_units": int(info.compute_units)}


def check(code: int) -> None:
    if code != TGH_OK:
        message = library().tgh_last_error().decode("utf-8", "replace")
        if code in (TGH_ERR_INVALID, TGH_ERR_SIZE):
            raise ValueError(message)
        raise RuntimeError(message)


def ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def default_config() -> Config:
    config = Config()
    library().tgh_config_default(C.byref(config))
    return config


class PackedAssembly:
    """Holds contiguous arrays alive for a tgh_assembly view."""

    def __init__(self, counts, placed, intrinsic, radius, min_bend_radius, origin, lengths, periodic,
                 pin_fiber_ends=False, vertex_pinned=None, formation_steps=None, formation_layers=None):
        self.counts = np.ascontiguousarray(counts, dtype=np.uint32)
        self.placed = np.ascontiguousarray(placed, dtype=np.float64).reshape(-1)
        self.intrinsic = np.ascontiguousarray(intrinsic, dtype=np.float64).reshape(-1)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.min_bend_radius = (
            None if min_bend_radius is None else np.ascontiguousarray(min_bend_radius, dtype=np.float64)
        )
        self.vertex_pinned = None if vertex_pinned is None else np.ascontiguousarray(vertex_pinned, dtype=np.uint8)
        V = int(self.counts.sum())
        if self.placed.size != 3 * V or self.intrinsic.size != 3 * V:
            raise ValueError("placed/intrinsic must hold 3 coordinates per vertex")
        if self.radius.size != self.counts.size:
            raise ValueError("one radius per fiber is required")
        if self.min_bend_radius is not None and self.min_bend_radius.size != self.counts.size:
            raise ValueError("one min_bend_radius per fiber is required")
        if self.vertex_pinned is not None and self.vertex_pinned.size != V:
            raise ValueError("one pinned flag per vertex is required")
        self.struct = Assembly()
        self.struct.fiber_count = int(self.counts.size)
        self.struct.vertex_count = V
        self.struct.fiber_vertex_counts = ptr(self.counts, C.c_uint32)
        self.struct.placed = ptr(self.placed, C.c_double)
        self.struct.intrinsic = ptr(self.intrinsic, C.c_double)
        self.struct.fiber_radius = ptr(self.radius, C.c_double)
        self.struct.fiber_min_bend_radius = (
            ptr(self.min_bend_radius, C.c_double) if self.min_bend_radius is not None else None
        )
        self.struct.vertex_pinned = ptr(self.vertex_pinned, C.c_uint8) if self.vertex_pinned is not None else None
        for k in range(3):
            self.struct.cell_origin[k] = float(origin[k])
            self.struct.cell_lengths[k] = float(lengths[k])
            self.struct.cell_periodic[k] = 1 if periodic[k] else 0
        self.struct.pin_fiber_ends = 1 if pin_fiber_ends else 0
        F = self.counts.size
        self.formation_steps = None if formation_steps is None else np.ascontiguousarray(formation_steps, np.uint32)
        self.formation_layers = None if formation_layers is None else np.ascontiguousarray(formation_layers, np.uint32)
        for arr, name in ((self.formation_steps, "formation_steps"), (self.formation_layers, "formation_layers")):
            if arr is not None and arr.size != F:
                raise ValueError(f"one {name} entry per fiber is required")
        self.struct.fiber_formation_step = (ptr(self.formation_steps, C.c_uint32)
                                            if self.formation_steps is not None else None)
        self.struct.fiber_formation_layer = (ptr(self.formation_layers, C.c_uint32)
                                             if self.formation_layers is not None else None)

    @property
    def vertex_count(self) -> int:
        return int(self.struct.vertex_count)


class PackedLayout:
    """PackedAssembly's f32 structure of arrays for the uniform round layout
    (crates/tangle_relax/src/device/packed.rs:47-141), the input of
    DeviceWorld::new; holds the arrays alive for a tgh_packed view.
    `from_assembly` packs a PackedAssembly (f64 fiber arrays) the way
    from_assembly_with_options does (packed.rs:150-320): f32 positions, rest
    lengths as f32 of the f64 intrinsic distance (packed.rs:1058-1063), the
    bend limit 1/r_min on internal vertices, pins."""

    def __init__(self, positions, intrinsic_positions, segment_vertices, segment_fibers, segment_radii,
                 segment_rest_lengths, fiber_vertex_spans, vertex_max_curvature, vertex_pinned, cell_lower,
                 cell_upper, cell_periodic, *, segment_active=None, fiber_formation_layers=None,
                 fiber_formation_steps=None, adaptive=False, has_ovals=False):
        f32, u32 = np.float32, np.uint32
        self.positions = np.ascontiguousarray(positions, dtype=f32).reshape(-1)
        self.intrinsic_positions = np.ascontiguousarray(intrinsic_positions, dtype=f32).reshape(-1)
        self.segment_vertices = np.ascontiguousarray(segment_vertices, dtype=u32).reshape(-1)
        self.segment_fibers = np.ascontiguousarray(segment_fibers, dtype=u32)
        self.segment_radii = np.ascontiguousarray(segment_radii, dtype=f32)
        self.segment_rest_lengths = np.ascontiguousarray(segment_rest_lengths, dtype=f32)
        self.fiber_vertex_spans = np.ascontiguousarray(fiber_vertex_spans, dtype=u32).reshape(-1)
        self.vertex_max_curvature = np.ascontiguousarray(vertex_max_curvature, dtype=f32)
        self.vertex_pinned = np.ascontiguousarray(vertex_pinned, dtype=u32)
        self.segment_active = None if segment_active is None else np.ascontiguousarray(segment_active, dtype=u32)
        self.fiber_formation_layers = (None if fiber_formation_layers is None
                                       else np.ascontiguousarray(fiber_formation_layers, dtype=u32))
        self.fiber_formation_steps = (None if fiber_formation_steps is None
                                      else np.ascontiguousarray(fiber_formation_steps, dtype=u32))
        V, S, F = self.positions.size // 3, self.segment_fibers.size, self.fiber_vertex_spans.size // 2
        sizes = {"intrinsic_positions": (self.intrinsic_positions, 3 * V),
                 "segment_vertices": (self.segment_vertices, 2 * S), "segment_radii": (self.segment_radii, S),
                 "segment_rest_lengths": (self.segment_rest_lengths, S),
                 "vertex_max_curvature": (self.vertex_max_curvature, V), "vertex_pinned": (self.vertex_pinned, V),
                 "segment_active": (self.segment_active, S),
                 "fiber_formation_layers": (self.fiber_formation_layers, F),
                 "fiber_formation_steps": (self.fiber_formation_steps, F)}
        for name, (arr, n) in sizes.items():
            if arr is not None and arr.size != n:
                raise ValueError(f"{name} must hold {n} values")
        st = PackedStruct()
        st.fiber_count, st.vertex_count, st.segment_count = F, V, S
        st.positions = ptr(self.positions, C.c_float)
        st.intrinsic_positions = ptr(self.intrinsic_positions, C.c_float)
        st.segment_vertices = ptr(self.segment_vertices, C.c_uint32)
        st.segment_fibers = ptr(self.segment_fibers, C.c_uint32)
        st.segment_radii = ptr(self.segment_radii, C.c_float)
        st.segment_rest_lengths = ptr(self.segment_rest_lengths, C.c_float)
        st.segment_active = None if self.segment_active is None else ptr(self.segment_active, C.c_uint32)
        st.fiber_vertex_spans = ptr(self.fiber_vertex_spans, C.c_uint32)
        st.fiber_formation_layers = (None if self.fiber_formation_layers is None
                                     else ptr(self.fiber_formation_layers, C.c_uint32))
        st.fiber_formation_steps = (None if self.fiber_formation_steps is None
                                    else ptr(self.fiber_formation_steps, C.c_uint32))
        st.vertex_max_curvature = ptr(self.vertex_max_curvature, C.c_float)
        st.vertex_pinned = ptr(self.vertex_pinned, C.c_uint32)
        for k in range(3):
            st.cell_lower[k] = float(cell_lower[k])
            st.cell_upper[k] = float(cell_upper[k])
            st.cell_periodic[k] = 1 if cell_periodic[k] else 0
        st.adaptive = 1 if adaptive else 0
        st.has_ovals = 1 if has_ovals else 0
        self.struct = st

    @property
    def vertex_count(self) -> int:
        return int(self.struct.vertex_count)

    @classmethod
    def from_assembly(cls, packed: PackedAssembly) -> "PackedLayout":
        counts = packed.counts.astype(np.int64)
        F, V = counts.size, int(counts.sum())
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
        intr64 = packed.intrinsic.reshape(-1, 3)
        seg_vtx, seg_fib, rest = [], [], []
        for f in range(F):
            first = np.arange(starts[f], starts[f] + counts[f] - 1)
            seg_vtx.append(np.stack([first, first + 1], axis=1))
            seg_fib.append(np.full(first.size, f))
            d = intr64[first + 1] - intr64[first]
            rest.append(np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))
        seg_vtx = np.concatenate(seg_vtx)
        seg_fib = np.concatenate(seg_fib)
        kmax = np.zeros(V, dtype=np.float32)
        if packed.min_bend_radius is not None:
            for f in range(F):
                r = packed.min_bend_radius[f]
                if r > 0.0 and np.isfinite(r):
                    kmax[starts[f] + 1:starts[f] + counts[f] - 1] = np.float32(1.0 / r)
        pinned = np.zeros(V, dtype=np.uint32)
        if packed.vertex_pinned is not None:
            pinned |= (packed.vertex_pinned != 0).astype(np.uint32)
        if packed.struct.pin_fiber_ends:
            pinned[starts] = 1
            pinned[starts + counts - 1] = 1
        origin = [packed.struct.cell_origin[k] for k in range(3)]
        upper = [packed.struct.cell_origin[k] + packed.struct.cell_lengths[k] for k in range(3)]
        return cls(packed.placed.astype(np.float32), packed.intrinsic.astype(np.float32), seg_vtx, seg_fib,
                   packed.radius.astype(np.float32)[seg_fib], np.concatenate(rest).astype(np.float32),
                   np.stack([starts, counts], axis=1), kmax, pinned, np.float32(origin), np.float32(upper),
                   [packed.struct.cell_periodic[k] for k in range(3)],
                   fiber_formation_layers=packed.formation_layers, fiber_formation_steps=packed.formation_steps)


class World:
    """One resident device world (DeviceFiberWorld on the HIP backend), or
    several specimens relaxed together (World.batch)."""

    def __init__(self, packed: PackedAssembly, config: Config):
        lib = library()
        self._packed = packed
        self._specimens = [packed]
        self._vertex_count = packed.vertex_count
        handle = C.c_void_p()
        check(lib.tgh_world_create(C.byref(packed.struct), C.byref(config), C.byref(handle)))
        self._handle = handle

    @classmethod
    def from_packed(cls, layout: "PackedLayout", config: Config) -> "World":
        """DeviceWorld::new from PackedAssembly's f32 arrays (tgh_world_create_packed)."""
        lib = library()
        self = cls.__new__(cls)
        self._packed = layout
        self._specimens = [layout]
        self._vertex_count = layout.vertex_count
        self._handle = None
        handle = C.c_void_p()
        check(lib.tgh_world_create_packed(C.byref(layout.struct), C.byref(config), C.byref(handle)))
        self._handle = handle
        return self

    @classmethod
    def batch(cls, specimens, config: Config) -> "World":
        """Independent specimens (PackedAssembly each) in one world."""
        specimens = list(specimens)
        if not specimens:
            raise ValueError("no specimens")
        lib = library()
        self = cls.__new__(cls)
        self._packed = specimens[0]
        self._specimens = specimens
        self._vertex_count = sum(p.vertex_count for p in specimens)
        arr = (Assembly * len(specimens))(*[p.struct for p in specimens])
        handle = C.c_void_p()
        check(lib.tgh_world_create_batch(arr, len(specimens), C.byref(config), C.byref(handle)))
        self._handle = handle
        return self

    @classmethod
    def restore(cls, packed: PackedAssembly, config: Config, checkpoint: "DeviceCheckpoint") -> "World":
        """A world for `packed` (the assembly the checkpointed world was made
        from) in the checkpoint's state (DeviceFiberWorld::restore,
        world.rs:692-760)."""
        lib = library()
        self = cls.__new__(cls)
        self._packed = packed
        self._specimens = [packed]
        self._vertex_count = packed.vertex_count
        c = checkpoint.struct()
        handle = C.c_void_p()
        check(lib.tgh_world_restore(C.byref(packed.struct), C.byref(config), C.byref(c), C.byref(handle)))
        self._handle = handle
        return self

    def checkpoint(self) -> "DeviceCheckpoint":
        """Downloads the persistent state (DeviceFiberWorld::checkpoint,
        world.rs:762-823)."""
        lib = library()
        sizes = CheckpointStruct()
        check(lib.tgh_world_checkpoint_sizes(self._handle, C.byref(sizes)))
        V, F = int(sizes.vertex_count), int(sizes.fiber_count)
        ck = DeviceCheckpoint(0, np.empty(3 * V, np.float32), np.empty(F, np.uint32), np.empty(3 * V, np.float32))
        if sizes.layer_target_count:
            ck.layer_target = dict(axis=0, first=0, last=0, stiffness=1.0, max_translation=1.0,
                                   targets=np.empty(sizes.layer_target_count, np.float32))
        if sizes.vertex_target_count:
            n = sizes.vertex_target_count
            ck.vertex_target = dict(axis=0, stiffness=1.0, max_translation=1.0, indices=np.empty(n, np.uint32),
                                    targets=np.empty(n, np.float32), commands=np.empty(n, np.float32))
        c = ck.struct()
        check(lib.tgh_world_checkpoint(self._handle, C.byref(c)))
        ck.total_iterations = int(c.total_iterations)
        ck.cell_lower = tuple(float(v) for v in c.cell_lower)
        ck.cell_upper = tuple(float(v) for v in c.cell_upper)
        if ck.layer_target is not None:
            ck.layer_target.update(axis=int(c.layer_axis), first=int(c.layer_first), last=int(c.layer_last),
                                   stiffness=float(c.layer_stiffness),
                                   max_translation=float(c.layer_max_translation))
        if ck.vertex_target is not None:
            ck.vertex_target.update(axis=int(c.vertex_axis), stiffness=float(c.vertex_stiffness),
                                    max_translation=float(c.vertex_max_translation))
        return ck

    @property
    def specimen_count(self) -> int:
        return len(self._specimens)

    def specimen_status(self):
        out = (BatchStatus * len(self._specimens))()
        check(library().tgh_world_specimen_status(self._handle, out, len(self._specimens)))
        return list(out)

    def specimen_positions(self):
        """Positions split per specimen (each 3 * its vertex count floats)."""
        flat = self.download_positions()
        out, start = [], 0
        for p in self._specimens:
            out.append(flat[start:start + 3 * p.vertex_count])
            start += 3 * p.vertex_count
        return out

    def close(self) -> None:
        if getattr(self, "_handle", None):
            library().tgh_world_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run_batch(self, config: Config, iterations: int) -> BatchStatus:
        status = BatchStatus()
        check(library().tgh_world_run_batch(self._handle, C.byref(config), int(iterations), C.byref(status)))
        return status

    def info(self) -> WorldInfo:
        info = WorldInfo()
        check(library().tgh_world_info_get(self._handle, C.byref(info)))
        return info

    def download_positions(self) -> np.ndarray:
        n = 3 * self._vertex_count
        out = np.empty(n, dtype=np.float32)
        check(library().tgh_world_download_positions(self._handle, ptr(out, C.c_float), n))
        return out

    def upload_positions(self, positions: np.ndarray) -> None:
        arr = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        check(library().tgh_world_upload_positions(self._handle, ptr(arr, C.c_float), arr.size))

    def tuning(self) -> Tuning:
        t = Tuning()
        check(library().tgh_world_get_tuning(self._handle, C.byref(t)))
        return t

    def set_tuning(self, **changes) -> None:
        t = self.tuning()
        for key, value in changes.items():
            setattr(t, key, value)
        check(library().tgh_world_set_tuning(self._handle, C.byref(t)))

    # -- formation (world.rs:1723-1927, 2139-2180) -----------------------
    def activate_formation_step(self, maximum_step: int):
        """Activates the fibers inserted at steps <= maximum_step; returns the
        (active segments, active vertices) counts."""
        segs, verts = C.c_uint64(), C.c_uint64()
        check(library().tgh_world_activate_formation_step(self._handle, int(maximum_step), C.byref(segs),
                                                           C.byref(verts)))
        return int(segs.value), int(verts.value)

    def set_layer_targets(self, axis, targets, first_layer=0, last_layer=None, stiffness=0.25,
                          max_translation=1.0e-5):
        arr = np.ascontiguousarray(targets, dtype=np.float32)
        last = arr.size - 1 if last_layer is None else int(last_layer)
        check(library().tgh_world_set_layer_targets(self._handle, int(axis), ptr(arr, C.c_float), arr.size,
                                                     int(first_layer), last, float(stiffness),
                                                     float(max_translation)))

    def clear_layer_targets(self):
        check(library().tgh_world_clear_layer_targets(self._handle))

    def set_vertex_targets(self, axis, vertices, displacement, stiffness=0.25, max_translation=1.0e-5):
        arr = np.ascontiguousarray(vertices, dtype=np.uint32)
        check(library().tgh_world_set_vertex_targets(self._handle, int(axis), ptr(arr, C.c_uint32), arr.size,
                                                      float(displacement), float(stiffness), float(max_translation)))

    def clear_vertex_targets(self):
        check(library().tgh_world_clear_vertex_targets(self._handle))

    def formation_target_error(self):
        """(layer error or None, vertex error or None)."""
        le, ve = C.c_float(), C.c_float()
        hl, hv = C.c_uint32(), C.c_uint32()
        check(library().tgh_world_formation_target_error(self._handle, C.byref(le), C.byref(ve), C.byref(hl),
                                                         C.byref(hv)))
        return (float(le.value) if hl.value else None, float(ve.value) if hv.value else None)

    def rebuild_lists(self) -> None:
        check(library().tgh_world_rebuild_lists(self._handle))

    # compaction and contact capture (world.rs:1931-2334)
    def cell_bounds(self):
        """(lower, upper) of the resident cell, f32 values (cell_bounds)."""
        lo, hi = (C.c_float * 3)(), (C.c_float * 3)()
        check(library().tgh_world_cell_bounds(self._handle, lo, hi))
        return tuple(float(v) for v in lo), tuple(float(v) for v in hi)

    def compact_cell(self, new_lower, new_upper, kinematics="moving_walls") -> None:
        """Changes the cell and moves the active material with it
        (compact_cell); kinematics "rigid_fiber_centers", "moving_walls" or
        "affine_vertices"."""
        lo = (C.c_float * 3)(*[float(v) for v in new_lower])
        hi = (C.c_float * 3)(*[float(v) for v in new_upper])
        check(library().tgh_world_compact_cell(self._handle, lo, hi, kinematics_code(kinematics)))

    def compaction_metrics(self, correction_fraction: float, model: CompactionEnergyModel | None = None):
        """Pressures and penalty energies of the current state
        (compaction_metrics)."""
        model = model or CompactionEnergyModel()
        m = CompactionMetricsStruct()
        check(library().tgh_world_compaction_metrics(self._handle, float(correction_fraction),
                                                     float(model.contact_stiffness), float(model.stretch_stiffness),
                                                     float(model.bending_stiffness), C.byref(m)))
        return CompactionMetrics(tuple(float(v) for v in m.pressure), float(m.contact_energy),
                                 float(m.stretch_energy), float(m.bending_energy), float(m.total_energy),
                                 tuple((float(m.face_pressure[k][0]), float(m.face_pressure[k][1])) for k in range(3)))

    def capture_contacts(self, maximum_surface_gap: float, capacity: int) -> ContactCapture:
        """Inter-fiber segment pairs within the surface gap, sorted
        (capture_contacts)."""
        out = (ContactCandidateStruct * max(1, int(capacity)))()
        count, overflow = C.c_size_t(), C.c_uint32()
        check(library().tgh_world_capture_contacts(self._handle, float(maximum_surface_gap), int(capacity), out,
                                                   C.byref(count), C.byref(overflow)))
        cands = [SegmentContactCandidate(int(c.first_segment), int(c.second_segment), float(c.first_coordinate),
                                         float(c.second_coordinate), float(c.surface_gap), float(c.crossing_angle))
                 for c in out[: count.value]]
        return ContactCapture(cands, bool(overflow.value))

    def capture_contact_arrays(self, maximum_surface_gap: float, capacity: int):
        """capture_contacts as arrays: (pairs (n, 2) u32, coordinates (n, 2),
        gaps (n,), angles (n,), overflow)."""
        out = np.empty((max(1, int(capacity)), 6), dtype=np.float32)
        count, overflow = C.c_size_t(), C.c_uint32()
        check(library().tgh_world_capture_contacts(self._handle, float(maximum_surface_gap), int(capacity),
                                                   out.ctypes.data_as(C.POINTER(ContactCandidateStruct)),
                                                   C.byref(count), C.byref(overflow)))
        n = int(count.value)
        return (out[:n, :2].view(np.uint32).copy(), out[:n, 2:4].copy(), out[:n, 4].copy(), out[:n, 5].copy(),
                bool(overflow.value))

    def set_fiber_maximum_curvatures(self, maximum_curvatures) -> None:
        """Per-fiber curvature limits, inverse length, 0 disables
        (set_fiber_maximum_curvatures)."""
        arr = np.ascontiguousarray(maximum_curvatures, dtype=np.float32)
        check(library().tgh_world_set_fiber_maximum_curvatures(self._handle, ptr(arr, C.c_float), arr.size))

    def cell_list(self):
        info = self.info()
        counts = np.empty(info.cell_count, dtype=np.uint32)
        offsets = np.empty(info.cell_count, dtype=np.uint32)
        segments = np.empty(info.segment_count, dtype=np.uint32)
        check(
            library().tgh_world_cell_list(
                self._handle, ptr(counts, C.c_uint32), ptr(offsets, C.c_uint32), ptr(segments, C.c_uint32),
                counts.size, segments.size,
            )
        )
        return counts, offsets, segments

    def neighbor_lists(self):
        info = self.info()
        counts = np.empty(info.segment_count, dtype=np.uint32)
        lists = np.zeros(max(1, info.segment_count * info.neighbor_block), dtype=np.uint32)
        check(
            library().tgh_world_neighbor_lists(self._handle, ptr(counts, C.c_uint32), ptr(lists, C.c_uint32), lists.size)
        )
        return counts, lists[: info.segment_count * info.neighbor_block].reshape(info.segment_count, -1)


def generate_population(spec: PopulationSpec):
    """Runs tgh_generate_population; returns (intrinsic, placed, radius,
    formation_layer), the last 0xffffffff for unlayered fibers."""
    nv = int(spec.segments_per_fiber) + 1
    n = int(spec.count)
    intrinsic = np.empty((n * nv, 3), dtype=np.float64)
    placed = np.empty((n * nv, 3), dtype=np.float64)
    radius = np.empty(n, dtype=np.float64)
    layers = np.empty(n, dtype=np.uint32)
    check(
        library().tgh_generate_population(
            C.byref(spec), ptr(intrinsic, C.c_double), ptr(placed, C.c_double), ptr(radius, C.c_double),
            ptr(layers, C.c_uint32), n * nv
        )
    )
    return intrinsic, placed, radius, layers
