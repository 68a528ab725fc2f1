"""Restart checkpoints for long recipe runs (crates/tangle_checkpoint/src/lib.rs,
crates/tangle_python/src/checkpoint.rs).

A checkpoint file holds the case identifier, the persistent device state of
the world (`_native.DeviceCheckpoint`: positions, fiber activity, wall
reactions, the lifetime iteration count, the held formation targets and the
cell, which compaction may have changed since creation) and
the recipe's progress (the operation in progress and where it started, the
held layer targets, the events so far). Scratch broad-phase state is not
kept; the first batch after a restore rebuilds it. Further named device
states may ride along (a compaction in progress keeps its last accepted
state this way); their arrays carry the name as a prefix.

File layout (little-endian), written to a temporary file and renamed over
the target so a reader never sees a partial checkpoint:

    8 bytes   magic b"TGHAMD01"
    4 bytes   schema (u32, currently 1)
    8 bytes   header length H (u64)
    H bytes   UTF-8 JSON header: case_id, scalars, recipe progress, and for
              every array its name, dtype, element count and byte offset
    ...       the arrays, array k at (start of this part + its offset); the
              offsets are multiples of 8

The reference stores its checkpoints with bincode under its own magic
(b"TANGLE01", schema 4) over Rust types that have no counterpart here; this
format keeps the same content for the device world and the recipe progress.
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._native import DeviceCheckpoint

MAGIC = b"TGHAMD01"
SCHEMA = 1
_DTYPES = {"f32": np.float32, "u32": np.uint32}


class CheckpointError(RuntimeError):
    """A checkpoint could not be read or does not fit the request
    (CheckpointError, tangle_checkpoint/src/lib.rs:114-175)."""


@dataclass
class CheckpointSettings:
    """Periodic restart checkpoint and optional resume settings
    (CheckpointSettings, crates/tangle_python/src/checkpoint.rs:10-100)."""

    case_id: str
    path: object
    interval_iterations: int = 500
    resume: bool = False
    resume_path: object = None
    resume_case_id: str | None = None
    fresh_formation_on_resume: bool = False

    def __post_init__(self):
        self.path = Path(self.path)
        if self.resume_path is not None:
            self.resume_path = Path(self.resume_path)
        self.validate()

    def validate(self) -> None:
        if not str(self.case_id).strip():
            raise ValueError("checkpoint case_id must not be empty")
        if int(self.interval_iterations) <= 0:
            raise ValueError("checkpoint interval_iterations must be positive")
        if self.resume_case_id is not None and not str(self.resume_case_id).strip():
            raise ValueError("resume_case_id must not be empty when supplied")

    @property
    def source(self) -> tuple[Path, str]:
        """(path, case id) to resume from."""
        return (self.resume_path or self.path, self.resume_case_id or self.case_id)


def _arrays(device: DeviceCheckpoint):
    out = {"positions": device.positions, "fiber_active": device.fiber_active,
           "wall_reactions": device.wall_reactions}
    if device.layer_target is not None:
        out["layer_targets"] = device.layer_target["targets"]
    if device.vertex_target is not None:
        vt = device.vertex_target
        out.update(vertex_indices=vt["indices"], vertex_targets=vt["targets"], vertex_commands=vt["commands"])
    return out


def _scalars(device: DeviceCheckpoint) -> dict:
    scalars = {"total_iterations": int(device.total_iterations)}
    if device.cell_lower is not None:
        scalars["cell"] = [[float(v) for v in device.cell_lower], [float(v) for v in device.cell_upper]]
    if device.layer_target is not None:
        scalars["layer_target"] = {k: v for k, v in device.layer_target.items() if k != "targets"}
    if device.vertex_target is not None:
        scalars["vertex_target"] = {k: v for k, v in device.vertex_target.items()
                                    if k not in ("indices", "targets", "commands")}
    return scalars


def _device(dev: dict, arrays: dict, prefix: str = "") -> DeviceCheckpoint:
    ck = DeviceCheckpoint(int(dev["total_iterations"]), arrays[prefix + "positions"],
                          arrays[prefix + "fiber_active"], arrays[prefix + "wall_reactions"])
    if "cell" in dev:
        ck.cell_lower, ck.cell_upper = tuple(dev["cell"][0]), tuple(dev["cell"][1])
    if "layer_target" in dev:
        ck.layer_target = dict(dev["layer_target"], targets=arrays[prefix + "layer_targets"])
    if "vertex_target" in dev:
        ck.vertex_target = dict(dev["vertex_target"], indices=arrays[prefix + "vertex_indices"],
                                targets=arrays[prefix + "vertex_targets"], commands=arrays[prefix + "vertex_commands"])
    for name in ("positions", "wall_reactions"):
        if arrays[prefix + name].size % 3:
            raise CheckpointError(f"checkpoint codec failed: {prefix}{name} is not 3 per vertex")
    return ck


def save(path, case_id: str, device: DeviceCheckpoint, recipe: dict, extras: dict | None = None) -> int:
    """Writes one checkpoint atomically; returns its size in bytes. `extras`
    maps names to further device states kept in the same file."""
    path = Path(path)
    arrays = _arrays(device)
    scalars = _scalars(device)
    extra_scalars = {}
    for name, extra in (extras or {}).items():
        extra_scalars[name] = _scalars(extra)
        arrays.update({f"{name}.{k}": v for k, v in _arrays(extra).items()})
    table, offset = [], 0
    for name, arr in arrays.items():
        dtype = "f32" if arr.dtype == np.float32 else "u32"
        table.append({"name": name, "dtype": dtype, "count": int(arr.size), "offset": offset})
        offset += -(-arr.nbytes // 8) * 8
    head = {"case_id": case_id, "device": scalars, "recipe": recipe, "arrays": table}
    if extra_scalars:
        head["extra_devices"] = extra_scalars
    header = json.dumps(head, separators=(",", ":")).encode()
    tmp = path.with_name(f".{path.name}.tmp")
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(np.uint32(SCHEMA).tobytes())
        f.write(np.uint64(len(header)).tobytes())
        f.write(header)
        base = f.tell()
        for entry, arr in zip(table, arrays.values()):
            f.seek(base + entry["offset"])
            f.write(np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder("<"), copy=False).tobytes())
        pad = (-f.tell()) % 8
        f.write(b"\0" * pad)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)
    return path.stat().st_size


def load(path, case_id: str) -> tuple[DeviceCheckpoint, dict]:
    """Reads a checkpoint; raises CheckpointError on a foreign or
    mismatched file (load_checkpoint, tangle_checkpoint/src/lib.rs:374-404)."""
    ck, recipe, _ = load_with_extras(path, case_id)
    return ck, recipe


def load_with_extras(path, case_id: str) -> tuple[DeviceCheckpoint, dict, dict]:
    """As `load`, and the further named device states the file holds."""
    path = Path(path)
    try:
        data = path.read_bytes()
    except OSError as err:
        raise CheckpointError(f"checkpoint I/O failed: {err}") from err
    if len(data) < 20 or data[:8] != MAGIC or int(np.frombuffer(data[8:12], "<u4")[0]) != SCHEMA:
        schema = int(np.frombuffer(data[8:12], "<u4")[0]) if len(data) >= 12 else -1
        raise CheckpointError(f"incompatible checkpoint magic {data[:8]!r} or schema {schema}")
    hlen = int(np.frombuffer(data[12:20], "<u8")[0])
    try:
        header = json.loads(data[20:20 + hlen].decode())
    except (UnicodeDecodeError, json.JSONDecodeError) as err:
        raise CheckpointError(f"checkpoint codec failed: {err}") from err
    if header.get("case_id") != case_id:
        raise CheckpointError(f"checkpoint case '{header.get('case_id')}' does not match requested case '{case_id}'")
    base = 20 + hlen
    arrays = {}
    for entry in header["arrays"]:
        dt = np.dtype(_DTYPES[entry["dtype"]]).newbyteorder("<")
        start = base + int(entry["offset"])
        stop = start + int(entry["count"]) * dt.itemsize
        if stop > len(data):
            raise CheckpointError("checkpoint codec failed: truncated array data")
        arrays[entry["name"]] = np.frombuffer(data[start:stop], dtype=dt).astype(dt.newbyteorder("="))
    try:
        ck = _device(header["device"], arrays)
        extras = {name: _device(dev, arrays, name + ".") for name, dev in header.get("extra_devices", {}).items()}
    except KeyError as err:
        raise CheckpointError(f"checkpoint codec failed: missing array {err}") from err
    return ck, header["recipe"], extras
