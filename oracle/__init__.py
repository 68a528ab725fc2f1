"""ORACLE — test infrastructure only.

CPU restatement of TANGLE's relaxation loop (relax_oracle.cpp, built into
liboracle.so by this directory's Makefile) and of its population generator
(population_ref.py). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this package, and only as the checker or as the timed
CPU baseline; the product (tangle_amd) never imports it.

OracleWorld mirrors tangle_amd._native.World so a test can drive both on the
same inputs and compare.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from tangle_amd._native import Assembly, BatchStatus, Config, PackedAssembly, WorldInfo

HERE = Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIBRARY_PATH


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            build()
        lib = C.CDLL(str(LIBRARY_PATH))
        u32p, f32p = C.POINTER(C.c_uint32), C.POINTER(C.c_float)
        sig = {
            "orc_world_create": (C.c_int, [C.POINTER(Assembly), C.POINTER(Config), C.POINTER(C.c_void_p)]),
            "orc_world_run_batch": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_uint32, C.POINTER(BatchStatus)]),
            "orc_world_download_positions": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
            "orc_world_upload_positions": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
            "orc_world_info": (C.c_int, [C.c_void_p, C.POINTER(WorldInfo)]),
            "orc_world_rebuild_now": (C.c_int, [C.c_void_p]),
            "orc_world_cell_list": (C.c_int, [C.c_void_p, u32p, u32p, u32p]),
            "orc_world_neighbor_lists": (C.c_int, [C.c_void_p, u32p, u32p]),
            "orc_world_contact_state": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
            "orc_world_activate": (None, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "orc_world_set_layer_targets": (None, [C.c_void_p, C.c_uint32, f32p, C.c_size_t, C.c_uint32, C.c_uint32,
                                                   C.c_float, C.c_float]),
            "orc_world_clear_layer_targets": (None, [C.c_void_p]),
            "orc_world_set_vertex_targets": (None, [C.c_void_p, C.c_uint32, u32p, C.c_size_t, C.c_float, C.c_float,
                                                    C.c_float]),
            "orc_world_clear_vertex_targets": (None, [C.c_void_p]),
            "orc_world_formation_target_error": (None, [C.c_void_p, f32p, f32p, C.POINTER(C.c_uint32),
                                                  C.POINTER(C.c_uint32)]),
            "orc_world_compact_cell": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                 C.c_uint32]),
            "orc_world_cell_bounds": (None, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
            "orc_world_compaction_metrics": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                                                    C.POINTER(C.c_float)]),
            "orc_world_capture_contacts": (C.c_int, [C.c_void_p, C.c_float, C.c_size_t, C.POINTER(C.c_float),
                                                     C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
            "orc_world_set_fiber_maximum_curvatures": (None, [C.c_void_p, C.POINTER(C.c_float)]),
            "orc_world_gs_fibers": (None, [C.c_void_p, u32p, u32p]),
            "orc_gs_counters": (None, [C.POINTER(C.c_uint64)]),
            "orc_world_destroy": (None, [C.c_void_p]),
            "orc_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(code: int) -> None:
    if code != 0:
        message = library().orc_last_error().decode()
        raise (ValueError if code in (-1, -3) else RuntimeError)(message)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleWorld:
    def __init__(self, packed: PackedAssembly, config: Config):
        self._packed = packed
        h = C.c_void_p()
        _check(library().orc_world_create(C.byref(packed.struct), C.byref(config), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            library().orc_world_destroy(self._h)
            self._h = None

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
        st = BatchStatus()
        _check(library().orc_world_run_batch(self._h, C.byref(config), int(iterations), C.byref(st)))
        return st

    def info(self) -> WorldInfo:
        info = WorldInfo()
        _check(library().orc_world_info(self._h, C.byref(info)))
        return info

    def download_positions(self) -> np.ndarray:
        n = 3 * self._packed.vertex_count
        out = np.empty(n, dtype=np.float32)
        _check(library().orc_world_download_positions(self._h, _p(out, C.c_float), n))
        return out

    def upload_positions(self, positions) -> None:
        arr = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        _check(library().orc_world_upload_positions(self._h, _p(arr, C.c_float), arr.size))

    def rebuild_lists(self) -> None:
        _check(library().orc_world_rebuild_now(self._h))

    # formation, as _native.World
    def activate_formation_step(self, maximum_step):
        s, v = C.c_uint64(), C.c_uint64()
        library().orc_world_activate(self._h, int(maximum_step), C.byref(s), C.byref(v))
        return int(s.value), int(v.value)

    def set_layer_targets(self, axis, targets, first_layer=0, last_layer=None, stiffness=0.25, max_translation=1.0e-5):
        arr = np.ascontiguousarray(targets, dtype=np.float32)
        last = arr.size - 1 if last_layer is None else int(last_layer)
        library().orc_world_set_layer_targets(self._h, int(axis), _p(arr, C.c_float), arr.size, int(first_layer), last,
                                              float(stiffness), float(max_translation))

    def clear_layer_targets(self):
        library().orc_world_clear_layer_targets(self._h)

    def set_vertex_targets(self, axis, vertices, displacement, stiffness=0.25, max_translation=1.0e-5):
        arr = np.ascontiguousarray(vertices, dtype=np.uint32)
        library().orc_world_set_vertex_targets(self._h, int(axis), _p(arr, C.c_uint32), arr.size, float(displacement),
                                               float(stiffness), float(max_translation))

    def clear_vertex_targets(self):
        library().orc_world_clear_vertex_targets(self._h)

    def formation_target_error(self):
        """(layer error or None, vertex error or None), as the reference's
        Option results (world.rs:1867-1927)."""
        le, ve = C.c_float(), C.c_float()
        hl, hv = C.c_uint32(), C.c_uint32()
        library().orc_world_formation_target_error(self._h, C.byref(le), C.byref(ve), C.byref(hl), C.byref(hv))
        return (float(le.value) if hl.value else None, float(ve.value) if hv.value else None)

    def cell_list(self):
        info = self.info()
        counts = np.empty(info.cell_count, dtype=np.uint32)
        offsets = np.empty(info.cell_count, dtype=np.uint32)
        segs = np.empty(info.segment_count, dtype=np.uint32)
        _check(library().orc_world_cell_list(self._h, _p(counts, C.c_uint32), _p(offsets, C.c_uint32),
                                             _p(segs, C.c_uint32)))
        return counts, offsets, segs

    def neighbor_lists(self):
        info = self.info()
        counts = np.empty(info.segment_count, dtype=np.uint32)
        lists = np.zeros(max(1, info.segment_count * info.neighbor_block), dtype=np.uint32)
        _check(library().orc_world_neighbor_lists(self._h, _p(counts, C.c_uint32), _p(lists, C.c_uint32)))
        return counts, lists[: info.segment_count * info.neighbor_block].reshape(info.segment_count, -1)

    def curvature_fiber_counts(self):
        """Per fiber, at its last curvature cleanup: (projections applied,
        triplets the GPU sweep visits)."""
        F = self.info().fiber_count
        projections = np.zeros(F, dtype=np.uint32)
        visits = np.zeros(F, dtype=np.uint32)
        library().orc_world_gs_fibers(self._h, _p(projections, C.c_uint32), _p(visits, C.c_uint32))
        return projections, visits

    def contact_state(self):
        info = self.info()
        corr = np.empty(6 * info.segment_count, dtype=np.float32)
        smax = np.empty(info.segment_count, dtype=np.float32)
        curv = np.empty(info.vertex_count, dtype=np.float32)
        _check(library().orc_world_contact_state(self._h, _p(corr, C.c_float), _p(smax, C.c_float),
                                                 _p(curv, C.c_float)))
        return corr.reshape(-1, 6), smax, curv

    # compaction and contact capture, as _native.World
    def cell_bounds(self):
        lo, hi = (C.c_float * 3)(), (C.c_float * 3)()
        library().orc_world_cell_bounds(self._h, lo, hi)
        return tuple(float(v) for v in lo), tuple(float(v) for v in hi)

    def compact_cell(self, new_lower, new_upper, kinematics: int) -> None:
        lo = (C.c_float * 3)(*[float(v) for v in new_lower])
        hi = (C.c_float * 3)(*[float(v) for v in new_upper])
        _check(library().orc_world_compact_cell(self._h, lo, hi, int(kinematics)))

    def compaction_metrics(self, correction_fraction, contact_stiffness=1.0, stretch_stiffness=1.0,
                           bending_stiffness=1.0) -> np.ndarray:
        """The 13 outputs of measure_compaction_metrics (kernels.rs:811-923)."""
        out = np.empty(13, dtype=np.float32)
        library().orc_world_compaction_metrics(self._h, float(correction_fraction), float(contact_stiffness),
                                               float(stretch_stiffness), float(bending_stiffness),
                                               _p(out, C.c_float))
        return out

    def capture_contact_arrays(self, maximum_surface_gap, capacity):
        out = np.zeros((max(1, int(capacity)), 6), dtype=np.float32)
        count, overflow = C.c_size_t(), C.c_uint32()
        _check(library().orc_world_capture_contacts(self._h, float(maximum_surface_gap), int(capacity),
                                                    _p(out, C.c_float), C.byref(count), C.byref(overflow)))
        n = int(count.value)
        return (out[:n, :2].view(np.uint32).copy(), out[:n, 2:4].copy(), out[:n, 4].copy(), out[:n, 5].copy(),
                bool(overflow.value))

    def set_fiber_maximum_curvatures(self, values) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float32)
        library().orc_world_set_fiber_maximum_curvatures(self._h, _p(arr, C.c_float))
