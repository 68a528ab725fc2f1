"""Writes the golden vectors in this directory from the oracle (run from the
repository root: python tests/golden/make_golden.py).

The vectors are regression pins of the oracle itself (the reference cannot be
run here); tests/test_golden.py checks the oracle against them on the CPU and
tests/test_gpu_golden.py checks the HIP library against them on the GPU.

  config1_golden.npz   40 straight fibers (d = 10 um, L = 300 um, 30
                       segments, 300 um periodic cube, seed 1): generated
                       positions, the cell list and neighbour lists at
                       upload, and positions and status after 25 iterations
                       at the hard-gate settings with force_full_iterations.
  bimodal_golden.npz   the bimodal felt at scale 0.02 (same fill, mix and
                       bend limits): the same quantities after 20 iterations.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import fixtures as fx  # noqa: E402
import oracle  # noqa: E402

CASES = {
    "config1": dict(recipe=lambda: fx.config1_recipe(), d=10e-6, iterations=25),
    "bimodal": dict(recipe=lambda: fx.bimodal_recipe(scale=0.02), d=7e-6, iterations=20),
}


def case_inputs(name):
    spec = CASES[name]
    packed = spec["recipe"]()._packed(False)
    cfg = fx.hard_gate_settings(spec["d"], force_full_iterations=True).to_config(spec["d"])
    return packed, cfg, spec["iterations"]


def compute(name):
    packed, cfg, iterations = case_inputs(name)
    with oracle.OracleWorld(packed, cfg) as w:
        w.rebuild_lists()
        counts, offsets, segments = w.cell_list()
        ncount, lists = w.neighbor_lists()
        status = w.run_batch(cfg, iterations)
        out = dict(
            placed=packed.placed.astype(np.float64),
            cell_counts=counts, cell_offsets=offsets, cell_segments=segments,
            neighbor_counts=ncount, neighbor_lists=lists,
            positions=w.download_positions(),
            total_iterations=np.array([status.total_iterations], dtype=np.uint64),
            max_penetration=np.array([status.max_penetration], dtype=np.float32),
            max_curvature_ratio=np.array([status.max_curvature_ratio], dtype=np.float32),
            iterations=np.array([iterations], dtype=np.uint32),
        )
    return out


def main():
    for name in CASES:
        np.savez_compressed(HERE / f"{name}_golden.npz", **compute(name))
        print("wrote", HERE / f"{name}_golden.npz")


if __name__ == "__main__":
    main()
