"""Benchmark: relaxation throughput and time to hard admissibility on one MI355X.

Workload: the bimodal felt specimen — 2,000 fibers 1 mm long in a 1 mm
periodic cube, 1,763 of 7 um and 237 of 19 um diameter (equal volume,
13.5 % fill), segment length ~2 d (131,335 segments), planar-biased
(Schladitz beta = 1/0.3 about z in the convention where beta > 1 lays fibers
flat), bend limits 200 um / 60 um, fp32 positions, seeded synthetic input.

A step is one relaxation batch of 128 iterations (iterations_per_batch)
starting from the generated (unrelaxed) specimen, run with
force_full_iterations so every iteration of the step executes. The timed
region is K steps; positions are reset to the generated specimen before it,
outside the timing. value = segment-iterations per second over the timed
region (segments x iterations / seconds).

Also reported: the wall time for Recipe.run to reach the hard admissibility
gates (max penetration <= 1e-3 d_min, max curvature ratio <= 1 + 1e-5),
plain and with the learned warm start on five specimens (training time
reported and counted separately), the oracle replaying the plain plan to the
same gates on the host cores, the HBM roofline of the dominant kernel from
HIP events captured in the timed batches and of the whole iteration's
algorithmic bytes, the same throughput legs on the slab and on 256 batched
specimens, and the CPU baseline (the C++ restatement of the reference loop,
OpenMP on the host cores) on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

METRIC = "segment-iters/s + time-to-hard-admissibility, learned vs plain, 1 MI355X"
PEAK_HBM_GBS = 8000.0
GROUPS = ["rebuild", "contact", "measure_assess", "vertex_passes", "curvature_gs"]


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed relaxation batches of 128 iterations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128, help="iterations per step")
    ap.add_argument("--tta-max-iterations", type=int, default=40000, help="budget of the final hard-gate stage")
    ap.add_argument("--tta-warm", type=int, default=2560, help="default-settings iterations before the final stage")
    ap.add_argument("--skip-tta", action="store_true", help="skip the time-to-admissibility run")
    ap.add_argument("--skip-learned", action="store_true", help="skip the learned warm-start run")
    ap.add_argument("--learned-warm", type=int, default=320, help="default-settings iterations of the learned plan")
    ap.add_argument("--learned-scale", type=float, default=8.0, help="proposal scale (intervals extrapolated)")
    ap.add_argument("--learned-interval", type=int, default=32, help="iterations between proposals")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iterations", type=int, default=384, help="iterations in the CPU baseline sample")
    ap.add_argument("--scale", type=float, default=1.0, help="specimen scale (1.0 = the 2,000-fiber workload)")
    ap.add_argument("--dump", type=str, default="", help="write final tta positions (.npy) here")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (reference kernel for the HBM counters)")
    ap.add_argument("--skip-batched", action="store_true", help="skip the batched-specimens throughput run")
    ap.add_argument("--skip-formation", action="store_true", help="skip the layered, needled felt workload")
    ap.add_argument("--skip-adjacent", action="store_true",
                    help="skip the compaction / contact-capture timings on the slab")
    ap.add_argument("--adjacent-only", action="store_true",
                    help="run only the compaction / contact-capture timings (for profiling)")
    ap.add_argument("--formation-plies", type=int, default=20, help="plies of the layered felt workload")
    ap.add_argument("--batched-specimens", type=int, default=256, help="specimens in the batched run")
    ap.add_argument("--batched-steps", type=int, default=2, help="timed 128-iteration batches of the batched run")
    ap.add_argument("--batched-tuning", type=str, default="",
                    help="skin,cell,contact_lanes,curvature_lanes for the batched run instead of the autotuner's "
                         "trials (for counter runs, whose serialised dispatches skew the trials' timings)")
    ap.add_argument("--curvature-lanes", type=int, default=-1,
                    help="lanes per fiber in the curvature sweep (0, 16, 32, 64; default: the world's choice)")
    ap.add_argument("--learned-specimens", type=int, default=5,
                    help="specimens of the learned-vs-plain comparison (the benchmark felt first)")
    ap.add_argument("--skip-cpu-tta", action="store_true",
                    help="skip the oracle's replay of the plain plan to the hard gates")
    ap.add_argument("--skip-slab", action="store_true", help="skip the slab throughput run")
    ap.add_argument("--workload", choices=["felt", "slab", "batched"], default="",
                    help="run only this workload's plain throughput leg (for profiling)")
    return ap.parse_args()


def curvature_bytes_per_launch(V: int) -> float:
    """Algorithmic bytes of one curvature-sweep launch: each vertex's position
    (16 B), static records vA and vB (32 B) and list snapshot (16 B) read once,
    its position written once (16 B)."""
    return 80.0 * V


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration); None when the summary does not cover the kernel."""
    path = ROOT / "profiles" / "hbm_counters_bimodal2000.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch ") and line.split()[1] == kernel:
            return int(line.split("=")[1])
    return None


def contact_bytes_per_launch(S: int, V: int, mean_list: float) -> float:
    """Algorithmic bytes of one contact launch: each cell-list slot's compact
    record (both endpoints and radius, 32 B) and static record (16 B:
    segment, fiber, radius, specimen) read once, its list count (4 B) and
    list entries (4 B each), and the two corrections written (32 B)."""
    return S * (32.0 + 16.0 + 4.0 + 4.0 * mean_list + 32.0)


def iteration_bytes(S: int, V: int, mean_list: float) -> float:
    """Algorithmic HBM bytes of one relaxation iteration (the whole loop, not
    one kernel): every vertex's position read and written and its intrinsic
    position, curvature limit, arc length and pin flag read once (48 B), and
    every segment's radius, fiber, rest length and list count read once
    (16 B) plus its list entries (4 B each); the list rebuild is excluded
    (amortised over the iterations between rebuilds, reported apart)."""
    return 48.0 * V + (16.0 + 4.0 * mean_list) * S


def whole_iteration_roofline(S, V, mean_list, iterations_per_second):
    nbytes = iteration_bytes(S, V, mean_list)
    achieved = nbytes * iterations_per_second / 1e9
    return {"bound": "hbm", "bytes_per_iteration": int(nbytes), "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5)}


def gs_kernel_name(lanes: int, walls: bool) -> str:
    """The curvature sweep's kernel as the profiler names it (8 lanes: the
    sequential walk, specialised on whether walls are tracked)."""
    seq = "tgh::k_curvature_gs_seq<true>" if walls else "tgh::k_curvature_gs_seq<false>"
    return {4: "tgh::k_curvature_gs_group<4>", 8: seq, 16: "tgh::k_curvature_gs_lean<16>",
            32: "tgh::k_curvature_gs_lean<32>", 64: "tgh::k_curvature_gs_wave64"}.get(lanes,
                                                                                      "tgh::k_curvature_gs_lane")


def host_cpu():
    """The host CPU as this process sees it: model, logical CPUs it may use
    and the physical cores under them (distinct package and core ids)."""
    allowed = sorted(os.sched_getaffinity(0))
    model, cores, cur = "unknown", set(), {}
    try:
        for line in open("/proc/cpuinfo").read().splitlines() + [""]:
            if not line.strip():
                if cur.get("processor") is not None and int(cur["processor"]) in allowed:
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                cur = {}
                continue
            key, _, value = line.partition(":")
            cur[key.strip()] = value.strip()
            if key.strip() == "model name":
                model = value.strip()
    except OSError:
        pass
    return {"model": model, "logical_cpus": len(allowed), "physical_cores": len(cores) or len(allowed)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of this "
                                   "build (the sharding plan is written up with the design)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import fixtures as fx
    import tangle_amd as tg
    from tangle_amd import _native

    if args.adjacent_only:
        print(json.dumps({"compaction_capture": adjacent_workload(fx, _native),
                          "checkpoint_export": checkpoint_export_workload(fx, tg, _native)}), flush=True)
        return

    d_min = 7e-6
    recipe = fx.bimodal_recipe(scale=args.scale)
    packed = recipe._packed(False)
    settings = fx.hard_gate_settings(d_min, force_full_iterations=True)
    cfg = settings.to_config(d_min)

    if args.workload:
        # one plain throughput leg only, for the profiler
        if args.workload == "felt":
            out = plain_leg(_native, packed, cfg, args.steps, args.warmup, args.batch, args.curvature_lanes)
        elif args.workload == "slab":
            out = slab_throughput(args, fx, _native)
        else:
            out = batched_throughput(args, fx, tg, _native)
        print(json.dumps({"workload": args.workload, "leg": out}, default=float), flush=True)
        return

    progress("felt throughput")
    leg = plain_leg(_native, packed, cfg, args.steps, args.warmup, args.batch, args.curvature_lanes)
    S, V = leg["segments"], leg["vertices"]
    lanes, contact_lanes, mean_list = leg["curvature_lanes"], leg["contact_lanes"], leg["mean_list_length"]
    total_iterations = leg["iterations"]
    group_ms, group_launches = leg["group_ms"], leg["group_launches"]
    value = leg["value"]
    ms_per_step = leg["ms_per_step"]

    # dominant kernel group by event time
    dominant = max(GROUPS, key=lambda g: group_ms[g])
    kernels = {
        "curvature_gs": (gs_kernel_name(lanes, not all(packed.struct.cell_periodic[k] for k in range(3))),
                         curvature_bytes_per_launch(V)),
        "contact": (f"tgh::k_contact<{contact_lanes}>", contact_bytes_per_launch(S, V, mean_list)),
    }

    def kernel_roofline(group):
        name, nbytes = kernels[group]
        avg_ms = group_ms[group] / max(1, group_launches[group])
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        return {
            "bound": "hbm",
            "kernel": name,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5),
            "traffic": profiled_traffic(name) if args.scale == 1.0 else None,
            "avg_launch_us": round(1000.0 * avg_ms, 2),
            "bytes_per_launch": int(nbytes),
        }

    roofline = kernel_roofline("curvature_gs")
    roofline["limiter"] = ("the sequential Gauss-Seidel projection chain of the slowest fiber "
                           "(one exact projection per step), not memory")
    roofline["contact"] = kernel_roofline("contact")
    roofline["whole_iteration"] = whole_iteration_roofline(S, V, mean_list, value / S)
    roofline["dominant_group"] = dominant
    roofline["group_ms_per_iteration"] = leg["group_ms_per_iteration"]
    roofline["timing_replay_bitwise_equal"] = leg["timing_replay_bitwise_equal"]

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "segment-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "bimodal felt: 2,000 fibers (1,763 x 7 um + 237 x 19 um), 1 mm in a 1 mm periodic cube, "
                        "planar-biased, plain HIP PBD",
            "segments": S,
            "vertices": V,
            "iterations_per_step": args.batch,
            "scale": args.scale,
            "cells": leg["cells"],
            "neighbor_block": leg["neighbor_block"],
            "mean_list_length": round(mean_list, 2),
            "rebuilds_in_timed_region": leg["rebuilds"],
            "curvature_lanes": lanes,
        },
        "roofline": roofline,
    }

    if not args.skip_tta:
        tta = {"plain": None, "learned": None, "gates": {"max_penetration_m": fx.HARD_PENETRATION,
                                                           "max_curvature_ratio": fx.HARD_CURVATURE_RATIO}}
        progress("plain time to the hard gates")
        plain, recorded, positions = plain_tta(args, fx, tg, _native, d_min, 0, record=True)
        tta["plain"] = plain
        if not args.skip_learned:
            progress("learned time to the hard gates")
            tta["learned"] = learned_tta(args, fx, tg, _native, d_min, plain)
            if plain.get("seconds") and tta["learned"].get("seconds"):
                tta["learned_speedup"] = round(plain["seconds"] / tta["learned"]["seconds"], 3)
        if not args.no_cpu and not args.skip_cpu_tta and plain.get("converged"):
            progress("oracle replay of the plain plan")
            tta["cpu_port"] = cpu_plan_replay(recorded, plain)
            if tta["cpu_port"].get("seconds") and tta.get("learned") and tta["learned"].get("seconds"):
                tta["learned_vs_cpu_port"] = round(tta["cpu_port"]["seconds"] / tta["learned"]["seconds"], 1)
                tta["plain_vs_cpu_port"] = round(tta["cpu_port"]["seconds"] / plain["seconds"], 1)
        result["time_to_hard_admissibility"] = tta
        if args.dump and positions is not None:
            np.save(args.dump, positions)

    if not args.skip_slab:
        progress("slab throughput")
        result["slab"] = slab_throughput(args, fx, _native)

    if not args.skip_batched:
        progress("batched throughput")
        result["batched"] = batched_throughput(args, fx, tg, _native)
        progress("batched time to the gate")
        result["batched_time_to_gate"] = batched_warm_start(fx, tg)

    if not args.skip_formation:
        progress("layered felt formation")
        result["formation"] = formation_workload(args, fx, tg)

    if not args.skip_adjacent:
        progress("compaction, capture, checkpoint")
        result["compaction_capture"] = adjacent_workload(fx, _native)
        result["checkpoint_export"] = checkpoint_export_workload(fx, tg, _native)

    if not args.no_cpu:
        progress("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(packed, cfg, S, args.cpu_iterations)

    print(json.dumps(result), flush=True)


def progress(stage: str) -> None:
    """A progress line on stderr (the result line goes to stdout)."""
    print(f"[bench] {stage}", file=sys.stderr, flush=True)


def plain_leg(_native, packed, cfg, steps, warmup, batch, curvature_lanes=-1):
    """K timed batches of `batch` iterations (force_full_iterations) from the
    generated positions, launched as captured graphs, then the same K
    batches replayed from the same start with every kernel group bracketed
    by HIP events on the world's stream (captured graphs do not expose
    event times, so those batches launch directly)."""
    import ctypes as C

    lib = _native.library()
    S = int(packed.vertex_count - packed.counts.size)
    V = int(packed.vertex_count)
    initial = packed.placed.astype(np.float32)
    world = _native.World(packed, cfg)
    if curvature_lanes >= 0:
        world.set_tuning(curvature_lanes=curvature_lanes)
    lanes = int(world.tuning().curvature_lanes)
    contact_lanes = int(world.tuning().contact_lanes)
    info = world.info()
    for _ in range(warmup):  # builds and caches the batch graph
        world.run_batch(cfg, batch)
    world.upload_positions(initial)
    rebuilds_before = world.info().neighbor_list_rebuilds
    t0 = time.perf_counter()
    total_iterations = 0
    for _ in range(steps):
        status = world.run_batch(cfg, batch)
        total_iterations += int(status.batch_iterations)
    seconds = time.perf_counter() - t0
    rebuilds = world.info().neighbor_list_rebuilds - rebuilds_before
    final_timed = world.download_positions()
    counts, _ = world.neighbor_lists()
    block = info.neighbor_block
    mean_list = float(np.mean(np.minimum(counts, block)))
    world.upload_positions(initial)
    _native.check(lib.tgh_world_set_timing(world._handle, 1))
    ms = (C.c_double * 5)()
    launches = (C.c_uint64 * 5)()
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    for _ in range(steps):
        world.run_batch(cfg, batch)
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    replay_matches = bool(np.array_equal(world.download_positions().view(np.uint32), final_timed.view(np.uint32)))
    world.close()
    group_ms = {GROUPS[g]: ms[g] for g in range(5)}
    group_launches = {GROUPS[g]: int(launches[g]) for g in range(5)}
    passes = total_iterations + steps  # each batch ends with a measuring pass
    return {
        "segments": S, "vertices": V, "value": round(S * total_iterations / seconds, 1), "unit": "segment-iters/s",
        "seconds": round(seconds, 4), "ms_per_step": round(1000.0 * seconds / steps, 3),
        "ms_per_iteration": round(1000.0 * seconds / total_iterations, 4), "iterations": total_iterations,
        "rebuilds": int(rebuilds), "mean_list_length": round(mean_list, 3), "neighbor_block": int(block),
        "cells": list(info.cells), "curvature_lanes": lanes, "contact_lanes": contact_lanes,
        "group_ms": group_ms, "group_launches": group_launches,
        "group_ms_per_iteration": {g: round(group_ms[g] / max(1, passes), 4) for g in GROUPS},
        "timing_replay_bitwise_equal": replay_matches,
    }


def slab_throughput(args, fx, _native):
    """The slab: 20,000 bimodal fibers 1 mm long in a 3 x 3 x 1 mm cell,
    periodic in x and y, walled in z (1.31 M segments), two timed 128-
    iteration batches from the generated state at the hard-gate settings,
    with the contact kernel's roofline and the whole iteration's."""
    packed = fx.slab_recipe()._packed(False)
    cfg = fx.hard_gate_settings(7e-6, force_full_iterations=True).to_config(7e-6)
    leg = plain_leg(_native, packed, cfg, 2, 1, args.batch)
    S, V, k = leg["segments"], leg["vertices"], leg["mean_list_length"]
    nbytes = contact_bytes_per_launch(S, V, k)
    avg_ms = leg["group_ms"]["contact"] / max(1, leg["group_launches"]["contact"])
    leg["workload"] = "slab: 20,000 bimodal fibers (17,630 x 7 um + 2,370 x 19 um), 3 x 3 x 1 mm, periodic x y, walled z"
    leg["contact_roofline"] = {"bound": "hbm", "kernel": f"tgh::k_contact<{leg['contact_lanes']}>",
                               "bytes_per_launch": int(nbytes), "avg_launch_us": round(1000.0 * avg_ms, 2),
                               "achieved": round(nbytes / (avg_ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                               "unit": "GB/s", "frac": round(nbytes / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    leg["whole_iteration"] = whole_iteration_roofline(S, V, k, leg["value"] / S)
    del leg["group_ms"], leg["group_launches"]
    return leg


class _BatchRecorder:
    """Records the world a Recipe.run creates and every batch it runs (a copy
    of the config and the iteration count), so that the oracle can replay
    the same plan."""

    def __init__(self, _native):
        self._native, self.created, self.calls = _native, [], []

    def __enter__(self):
        World = self._native.World
        self._init, self._run = World.__init__, World.run_batch
        rec = self

        def init(world, packed, config):
            rec.created.append((packed, type(config).from_buffer_copy(config)))
            rec._init(world, packed, config)

        def run(world, config, iterations):
            rec.calls.append((type(config).from_buffer_copy(config), int(iterations)))
            return rec._run(world, config, iterations)

        World.__init__, World.run_batch = init, run
        return self

    def __exit__(self, *exc):
        self._native.World.__init__, self._native.World.run_batch = self._init, self._run


def plain_tta(args, fx, tg, _native, d_min, seed_offset, record=False):
    """Plain HIP PBD plan to the hard gates: default relaxation, then a final
    SolvePolicy at the hard limits under the reference's contact_cleanup
    preset (the best of the schedules tried), through Recipe.run."""
    r = fx.staged_hard_gate_recipe(fx.bimodal_recipe(scale=args.scale, seed_offset=seed_offset),
                                   warm_iterations=args.tta_warm, cleanup_iterations=args.tta_max_iterations)
    settings = fx.hard_gate_settings(d_min, max_iterations=args.tta_warm + args.tta_max_iterations + 1)
    recorder = _BatchRecorder(_native)
    t0 = time.perf_counter()
    try:
        if record:
            with recorder:
                res = r.run(settings)
        else:
            res = r.run(settings)
        wall = time.perf_counter() - t0
        out = {"seconds": round(wall, 3), "iterations": res.iterations, "converged": res.converged,
               "max_penetration": res.max_penetration, "max_curvature_ratio": res.max_curvature_ratio,
               "plan": r.operations()[2:]}
        positions = res.positions
    except tg.RecipeError as err:
        out = {"seconds": None, "iterations": err.iteration, "converged": False, "error": str(err),
               "elapsed_s": round(time.perf_counter() - t0, 3)}
        positions = None
    return out, recorder, positions


def cpu_plan_replay(recorded, plain):
    """The oracle (the C++ restatement of the reference loop, OpenMP on the
    host cores) replays the plain plan's batches, config for config, from
    the same generated specimen: the CPU port's time to the hard gates. Its
    trajectory is the plain HIP path's bit for bit, so the same batches
    bring it to the gates; the final status is checked against them."""
    import oracle

    if len(recorded.created) != 1:
        return {"error": f"the plan created {len(recorded.created)} worlds; only one is replayed"}
    packed, config = recorded.created[0]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    with oracle.OracleWorld(packed, config) as w:
        t0 = time.perf_counter()
        status = None
        for cfg, iterations in recorded.calls:
            status = w.run_batch(cfg, iterations)
        seconds = time.perf_counter() - t0
    return {"seconds": round(seconds, 3), "iterations": int(status.total_iterations), "batches": len(recorded.calls),
            "converged": bool(status.converged), "max_penetration": status.max_penetration,
            "max_curvature_ratio": status.max_curvature_ratio,
            "same_iterations_as_plain": int(status.total_iterations) == int(plain["iterations"]),
            "kind": "port", "threads": threads, "host": host_cpu()}


def learned_tta(args, fx, tg, _native, d_min, plain0):
    """The learned warm start: a model trained on the GPU from plain-relaxer
    trajectories of two other specimens of the same distribution (seed
    offsets 7 and 13), then Recipe.run on the benchmark specimen with
    proposals every `--learned-interval` iterations, a shorter default-
    settings stage and the same final hard-gate stage as the plain plan.
    The same model and plans then run on further specimens (seed offsets
    101, 202, ...); speed-ups are reported per specimen, as minimum and
    geometric mean, and with the training time counted."""
    from tangle_amd import learned

    t0 = time.perf_counter()
    specimens = [fx.bimodal_recipe(scale=args.scale, seed_offset=o)._packed(False) for o in (7, 13)]
    model, losses = learned.train_warm_start(specimens, d_min, interval=args.learned_interval)
    rows = model.dataset_rows
    train_s = time.perf_counter() - t0
    # matrix-core kernels of the network, at the specimen's row count
    S = sum(p.vertex_count - p.counts.size for p in specimens[:1])
    fwd_ms, grad_ms = model.time_kernels(S, 20)
    tile_flops = 2.0 * (64 * 64 + 64 * 64 + 16 * 64)  # per row, as the 16x16x16 tiles compute it
    mlp = {"rows": S, "forward_us": round(1000 * fwd_ms, 2), "gradient_us": round(1000 * grad_ms, 2),
           "roofline": {"bound": "mfma", "kernel": "tgh::k_mlp_forward", "dtype": "bf16 (f32 accumulate)",
                        "achieved": round(tile_flops * S / (fwd_ms * 1e-3) / 1e12, 2), "peak": 2500.0,
                        "unit": "TFLOP/s",
                        "frac": round(tile_flops * S / (fwd_ms * 1e-3) / 1e12 / 2500.0, 4),
                        "input_GBps": round(128.0 * S / (fwd_ms * 1e-3) / 1e9, 1)}}

    def run_learned(seed_offset):
        ws = learned.WarmStart(model, interval=args.learned_interval, scale=args.learned_scale, settle=4,
                               active_iterations=10 ** 9)
        r = fx.staged_hard_gate_recipe(fx.bimodal_recipe(scale=args.scale, seed_offset=seed_offset),
                                       warm_iterations=args.learned_warm, cleanup_iterations=args.tta_max_iterations)
        settings = fx.hard_gate_settings(d_min, max_iterations=args.learned_warm + args.tta_max_iterations + 1)
        t1 = time.perf_counter()
        try:
            res = r.run(settings, warm_start=ws)
            return {"seconds": round(time.perf_counter() - t1, 3), "iterations": res.iterations,
                    "converged": res.converged, "max_penetration": res.max_penetration,
                    "max_curvature_ratio": res.max_curvature_ratio, "plan": r.operations()[2:],
                    "events": res.events[2:]}
        except tg.RecipeError as err:
            return {"seconds": None, "iterations": err.iteration, "converged": False, "error": str(err),
                    "elapsed_s": round(time.perf_counter() - t1, 3)}

    out = {"training_seconds": round(train_s, 3), "training_rows": int(rows),
           "training_loss": [round(float(losses[0]), 5), round(float(losses[-1]), 5)],
           "interval": args.learned_interval, "scale": args.learned_scale, "mlp": mlp}
    out.update(run_learned(0))
    specimens_out = []
    offsets = [0] + [101 * i for i in range(1, max(1, args.learned_specimens))]
    for o in offsets:
        progress(f"learned vs plain, specimen {o}")
        plain = plain0 if o == 0 else plain_tta(args, fx, tg, _native, d_min, o)[0]
        learned_run = out if o == 0 else run_learned(o)
        row = {"seed_offset": o, "plain_seconds": plain.get("seconds"), "plain_iterations": plain.get("iterations"),
               "learned_seconds": learned_run.get("seconds"), "learned_iterations": learned_run.get("iterations"),
               "both_converged": bool(plain.get("converged") and learned_run.get("converged"))}
        if row["both_converged"]:
            row["speedup"] = round(row["plain_seconds"] / row["learned_seconds"], 3)
        specimens_out.append(row)
    model.close()
    speedups = [r["speedup"] for r in specimens_out if "speedup" in r]
    summary = {"specimens": specimens_out, "converged_pairs": len(speedups)}
    if speedups:
        summary["min_speedup"] = min(speedups)
        summary["geomean_speedup"] = round(float(np.exp(np.mean(np.log(speedups)))), 3)
    if len(speedups) == len(specimens_out):
        plain_total = sum(r["plain_seconds"] for r in specimens_out)
        learned_total = sum(r["learned_seconds"] for r in specimens_out)
        n = len(specimens_out)
        # the model is trained once and reused: its cost is counted against
        # one specimen and against all of them
        summary["training_inclusive"] = {
            "one_specimen_speedup": round(specimens_out[0]["plain_seconds"] /
                                          (train_s + specimens_out[0]["learned_seconds"]), 3),
            f"{n}_specimens_speedup": round(plain_total / (train_s + learned_total), 3),
            "break_even_specimens": round(train_s / max(1e-9, (plain_total - learned_total) / n), 2)}
    out["five_specimens" if len(specimens_out) == 5 else "specimens"] = summary
    return out


def batched_throughput(args, fx, tg, _native):
    """Independent specimens in one batched world: fiber counts log-uniform
    in [50, 5,000], 10 um x 0.5 mm fibers of 25 segments, isotropic, each
    cube at 10 % fill. Every kernel launch serves all specimens; each step
    is a 128-iteration batch with force_full_iterations from the generated
    state. Reports segment-iterations/s and the event-timed kernel groups."""
    import ctypes as C

    recipes = fx.batched_specimen_recipes(count=args.batched_specimens)
    settings = tg.RelaxationSettings(force_full_iterations=True)
    packed = [r._packed(False) for r in recipes]
    cfg = settings.to_config(10e-6)
    S = sum(p.vertex_count - p.counts.size for p in packed)
    V = sum(p.vertex_count for p in packed)
    lib = _native.library()
    from tangle_amd.autotune import autotune
    t0 = time.perf_counter()
    if args.batched_tuning:
        from tangle_amd.autotune import Tuning
        skin, cell, cl, gl = args.batched_tuning.split(",")
        tuning = Tuning(cell_size_scale=float(cell), neighbor_skin_scale=float(skin), contact_lanes=int(cl),
                        curvature_lanes=int(gl))
    else:
        tuning = autotune(packed, cfg, iterations=32, cache=False)  # timed: the trials always run here
    tune_s = time.perf_counter() - t0
    default_tuning = {"cell_size_scale": cfg.cell_size_scale, "neighbor_skin_scale": cfg.neighbor_skin_scale}
    cfg = tuning.configure(cfg)
    t0 = time.perf_counter()
    world = _native.World.batch(packed, cfg)
    tuning.apply(world)
    create_s = time.perf_counter() - t0
    initial = world.download_positions()
    world.run_batch(cfg, args.batch)  # warmup: captures the batch graph
    world.upload_positions(initial)
    t0 = time.perf_counter()
    iterations = 0
    for _ in range(args.batched_steps):
        iterations += int(world.run_batch(cfg, args.batch).batch_iterations)
    seconds = time.perf_counter() - t0
    counts, _ = world.neighbor_lists()
    info = world.info()
    mean_list = float(np.mean(np.minimum(counts, info.neighbor_block)))
    world.upload_positions(initial)
    _native.check(lib.tgh_world_set_timing(world._handle, 1))
    ms = (C.c_double * 5)()
    launches = (C.c_uint64 * 5)()
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    for _ in range(args.batched_steps):
        world.run_batch(cfg, args.batch)
    _native.check(lib.tgh_world_timing(world._handle, ms, launches, 5, 1))
    world.close()
    group_ms = {GROUPS[g]: ms[g] for g in range(5)}
    contact_avg = group_ms["contact"] / max(1, launches[1])
    contact_bytes = contact_bytes_per_launch(S, V, mean_list)
    per_iter = {g: round(group_ms[g] / max(1, iterations + args.batched_steps), 4) for g in GROUPS}
    return {
        "workload": f"{len(packed)} independent specimens, fiber counts log-uniform in [50, 5000], 10 um x 0.5 mm, "
                    "25 segments, isotropic, 10 % fill each",
        "specimens": len(packed), "fibers": int(sum(p.counts.size for p in packed)), "segments": int(S),
        "vertices": int(V), "value": round(S * iterations / seconds, 1), "unit": "segment-iters/s",
        "ms_per_iteration": round(1000.0 * seconds / iterations, 3), "iterations": iterations,
        "create_seconds": round(create_s, 3), "mean_list_length": round(mean_list, 2),
        "autotune": {"seconds": round(tune_s, 2), "default": default_tuning,
                     "chosen": {"cell_size_scale": tuning.cell_size_scale,
                                "neighbor_skin_scale": tuning.neighbor_skin_scale,
                                "contact_lanes": tuning.contact_lanes, "curvature_lanes": tuning.curvature_lanes,
                                "iterations_per_rebuild": round(tuning.rebuild_interval, 2)},
                     "trial_ms": {k: {str(kk): round(vv, 3) for kk, vv in v.items()}
                                  for k, v in tuning.trials.items()}},
        "group_ms_per_iteration": per_iter,
        "whole_iteration": whole_iteration_roofline(S, V, mean_list, iterations / seconds),
        "contact_roofline": {"bound": "hbm", "kernel": f"tgh::k_contact<{tuning.contact_lanes}>", "bytes_per_launch": int(contact_bytes),
                             "avg_launch_us": round(1000.0 * contact_avg, 2),
                             "achieved": round(contact_bytes / (contact_avg * 1e-3) / 1e9, 1),
                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(contact_bytes / (contact_avg * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
    }


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of known size: its FETCH_SIZE and
    WRITE_SIZE counters per dispatch give the scale between those counters
    and bytes moved."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "dispatches": 5}), flush=True)


def batched_warm_start(fx, tg):
    """The learned warm start on a batched world: 64 of the batched
    specimens (fiber counts log-uniform in [50, 5,000], same seeds), relaxed
    together until every one has penetration <= 1e-3 d (d = 10 um), plain
    and with proposals from a model trained on two other specimens of the
    same kind (training time reported apart)."""
    from tangle_amd import learned

    d = 10e-6
    t0 = time.perf_counter()
    train = [r._packed(False) for r in fx.batched_specimen_recipes(count=2, n_min=1500, n_max=2000, seed=777)]
    model, _ = learned.train_warm_start(train, d, interval=32, default_iterations=1024, cleanup_iterations=0,
                                        epochs=12, batch_rows=16384)
    out = {"workload": "64 batched specimens (fiber counts log-uniform in [50, 5000], 10 um x 0.5 mm, 10 % fill) "
                       "to penetration <= 1e-3 d", "training_seconds": round(time.perf_counter() - t0, 3)}
    for label in ("plain", "learned"):
        recipes = fx.batched_specimen_recipes(count=64)
        settings = tg.RelaxationSettings(penetration_tolerance=1e-3 * d, curvature_ratio_tolerance=1e-5)
        ws = (learned.WarmStart(model, interval=32, scale=8.0, settle=4, active_iterations=10 ** 9)
              if label == "learned" else None)
        t1 = time.perf_counter()
        res = tg.relax_batch(recipes, settings, max_iterations=30000, warm_start=ws)
        out[label] = {"seconds": round(time.perf_counter() - t1, 3),
                      "slowest_specimen_iterations": max(r.iterations for r in res),
                      "converged": sum(r.converged for r in res)}
    model.close()
    if out["plain"]["converged"] == out["learned"]["converged"] == 64:
        out["learned_speedup"] = round(out["plain"]["seconds"] / out["learned"]["seconds"], 3)
        out["learned_speedup_training_included"] = round(
            out["plain"]["seconds"] / (out["learned"]["seconds"] + out["training_seconds"]), 3)
    return out


def adjacent_workload(fx, _native):
    """The compaction and contact-capture calls on the slab workload (20,000
    bimodal fibers, 1.31 M segments, walled in z) after 256 iterations:
    wall time per call (synchronised, so kernel time plus launch and copy
    overhead) and the HBM rate its algorithmic bytes imply. Algorithmic
    bytes per call: compaction metrics read 100 B per segment (record,
    activity, both corrections, both endpoints, rest length) and 20 B per
    vertex (ratio, wall reaction); moving walls and affine read 48 B and
    write 32 B per vertex; rigid centres read each vertex twice (16 B each)
    and write 32 B, plus 32 B per fiber."""
    pk = fx.slab_recipe()._packed(False)
    F, V = int(pk.counts.size), int(pk.vertex_count)
    S = V - F
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    cfg.force_full_iterations = 1
    out = {"workload": "slab: 20,000 bimodal fibers, 3 x 3 x 1 mm, periodic x y, walled z, after 256 iterations",
           "segments": S, "vertices": V}

    def timed(fn, repeats):
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            r = fn()
            times.append(time.perf_counter() - t0)
        return min(times), r

    def rate(nbytes, seconds):
        return {"ms": round(1e3 * seconds, 3), "algorithmic_MB": round(nbytes / 1e6, 2),
                "achieved_GBps": round(nbytes / seconds / 1e9, 1), "peak_GBps": PEAK_HBM_GBS,
                "frac": round(nbytes / seconds / 1e9 / PEAK_HBM_GBS, 4)}

    with _native.World(pk, cfg) as world:
        world.run_batch(cfg, 256)
        model = _native.CompactionEnergyModel(1.0, 1.0, 1.0)
        sec, m = timed(lambda: world.compaction_metrics(cfg.correction_fraction, model), 3)
        out["compaction_metrics"] = dict(rate(100 * S + 20 * V, sec),
                                         bound="latency: each sum is one lane adding its terms in index "
                                               "order, as the reference's single thread does",
                                         total_energy=m.total_energy)
        sec, cap = timed(lambda: world.capture_contact_arrays(1.0e-6, 1 << 22), 3)
        out["capture_contacts"] = {"ms": round(1e3 * sec, 3), "candidates": int(len(cap[2])),
                                   "overflow": bool(cap[4]),
                                   "note": "whole-segment cell list on the capture grid, 27-cell scan, "
                                           "device radix sort of the candidates, download"}
        lo, hi = world.cell_bounds()
        for name, nbytes in (("moving_walls", 80 * V), ("affine_vertices", 80 * V),
                             ("rigid_fiber_centers", 64 * V + 32 * F)):
            lo, hi = world.cell_bounds()
            dz = 0.002 * (hi[2] - lo[2])
            t0 = time.perf_counter()
            world.compact_cell([lo[0], lo[1], lo[2] + dz], [hi[0], hi[1], hi[2] - dz], name)
            out[f"compact_{name}"] = rate(nbytes, time.perf_counter() - t0)
    return out


def checkpoint_export_workload(fx, tg, _native):
    """Checkpoint and OVITO export of the benchmark felt after 256
    iterations: device-state download, atomic file write and read, restore
    into a new world, and one fiber-segment dump frame."""
    import tempfile
    from tangle_amd import checkpoint as ckpt

    recipe = fx.bimodal_recipe()
    pk = recipe._packed(False)
    cfg = fx.hard_gate_settings(7e-6).to_config(7e-6)
    cfg.force_full_iterations = 1
    out = {"workload": "bimodal felt (2,000 fibers, 133,335 vertices) after 256 iterations"}
    with _native.World(pk, cfg) as world, tempfile.TemporaryDirectory() as tmp:
        world.run_batch(cfg, 256)
        t0 = time.perf_counter()
        state = world.checkpoint()
        t1 = time.perf_counter()
        nbytes = ckpt.save(f"{tmp}/felt.ckpt", "bench", state, {"operation": 0})
        t2 = time.perf_counter()
        loaded, _ = ckpt.load(f"{tmp}/felt.ckpt", "bench")
        t3 = time.perf_counter()
        with _native.World.restore(pk, cfg, loaded) as again:
            t4 = time.perf_counter()
            same = bool(np.array_equal(again.download_positions().view(np.uint32),
                                       world.download_positions().view(np.uint32)))
        positions = world.download_positions()
        result = tg.RunResult(fiber_count=int(pk.counts.size), iterations=256, converged=False, max_penetration=0.0,
                              max_curvature_ratio=0.0, max_displacement=0.0, active_segments=0, active_vertices=0,
                              cell_count=0, uploaded_bytes=0, downloaded_bytes=0, events=[], warnings=[],
                              neighbor_list_rebuilds=0, wall_seconds=0.0, positions=positions,
                              vertex_counts=pk.counts.copy(), _fibers=list(recipe._fibers), _cell=recipe._cell)
        t5 = time.perf_counter()
        result.write_ovito(f"{tmp}/felt.dump")
        t6 = time.perf_counter()
        dump_bytes = os.path.getsize(f"{tmp}/felt.dump")
    out.update({"checkpoint_download_ms": round(1e3 * (t1 - t0), 2), "file_bytes": int(nbytes),
                "file_write_ms": round(1e3 * (t2 - t1), 2), "file_read_ms": round(1e3 * (t3 - t2), 2),
                "restore_ms": round(1e3 * (t4 - t3), 2), "restored_positions_bitwise_equal": same,
                "ovito_frame_ms": round(1e3 * (t6 - t5), 1), "ovito_frame_bytes": int(dump_bytes)})
    return out


def formation_workload(args, fx, tg):
    """The layered, needled felt: plies of 2,201 x 7 um + 299 x 19 um fibers
    on a 2 x 2 mm periodic footprint, inserted one formation step at a time
    and relaxed, 5 % of every ply's 19 um fibers needled 30 um down and
    released, then a contact and curvature cleanup (tests/fixtures.py
    layered_felt_recipe), through Recipe.run on one resident world. Reports
    active-segment iterations per second over the whole run, and the oracle
    on the fully stacked felt for a few iterations beside it."""
    import oracle

    recipe, settings = fx.layered_felt_recipe(plies=args.formation_plies)
    packed = recipe._packed(False)
    S = packed.vertex_count - packed.counts.size
    t0 = time.perf_counter()
    try:
        res = recipe.run(settings)
    except tg.RecipeError as err:
        return {"error": str(err), "seconds": round(time.perf_counter() - t0, 3)}
    wall = time.perf_counter() - t0
    out = {
        "workload": f"layered needled felt: {args.formation_plies} plies x (2,201 x 7 um + 299 x 19 um), "
                    f"1 mm fibers, 2 x 2 mm periodic footprint, bounded z",
        "fibers": len(recipe._fibers), "segments": int(S), "operations": len(recipe.operations()),
        "iterations": res.iterations, "post_recipe_iterations": res.post_recipe_iterations,
        "seconds": round(wall, 3), "value": round(res.segment_iterations / wall, 1), "unit": "segment-iters/s",
        "needled_vertices": int(sum(n[1] for n in res.needling)), "converged": res.converged,
        "max_penetration": res.max_penetration, "max_curvature_ratio": res.max_curvature_ratio,
        "warnings": res.warnings[:2],
    }
    if not args.no_cpu:
        cfg = settings.to_config(7e-6)
        with oracle.OracleWorld(packed, cfg) as w:
            t1 = time.perf_counter()
            status = w.run_batch(cfg, 4)
            t2 = time.perf_counter()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        out["cpu_baseline"] = {"value": round(S * int(status.batch_iterations) / (t2 - t1), 1),
                               "unit": "segment-iters/s", "cores": threads, "kind": "port",
                               "sample": f"{int(status.batch_iterations)} iterations of the fully stacked felt "
                                         f"from the generated positions ({t2 - t1:.1f} s)"}
    return out


def cpu_baseline(packed, cfg, S, iterations):
    """The oracle (C++ restatement of the reference loop, OpenMP) on the first
    `iterations` iterations of the same specimen, on the host cores."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    with oracle.OracleWorld(packed, cfg) as w:
        t1 = time.perf_counter()
        status = w.run_batch(cfg, iterations)
        t2 = time.perf_counter()
    done = int(status.batch_iterations)
    # the plain HIP path on the same iterations from the same state (graph
    # capture included), for a like-for-like ratio
    from tangle_amd import _native
    with _native.World(packed, cfg) as g:
        g.run_batch(cfg, 1)  # library and stream warm-up on a throwaway world
    with _native.World(packed, cfg) as g:
        t3 = time.perf_counter()
        gpu_status = g.run_batch(cfg, iterations)
        t4 = time.perf_counter()
    same = int(gpu_status.batch_iterations) == done
    return {
        "value": round(S * done / (t2 - t1), 1),
        "unit": "segment-iters/s",
        "cores": threads,
        "host": host_cpu(),
        "build": "g++ -O3 -march=x86-64-v3 -fopenmp (oracle/Makefile)",
        "kind": "port",
        "sample": f"first {done} iterations of the same specimen from the generated state "
                  f"({t2 - t1:.1f} s; upload and sizing {t1 - t0:.2f} s excluded)",
        "gpu_same_sample_seconds": round(t4 - t3, 4),
        "gpu_vs_cpu_same_sample": round((t2 - t1) / (t4 - t3), 1) if same else None,
    }


if __name__ == "__main__":
    main()
