"""The HIP library against the committed golden vectors, through the C ABI:
cell list and neighbour lists bit-exact at upload; positions after the fixed
iterations within 1e-4 d and bit-exact, for the straight-fiber case and for
the bimodal felt with its bend limits (the curvature sweep's acosf is the
host libm's algorithm)."""

from __future__ import annotations

import numpy as np
import pytest

from golden import make_golden
from tangle_amd import _native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(make_golden.CASES))
def test_hip_matches_golden(hip_device, name):
    print(f"[case] golden {name}", flush=True)
    with np.load(make_golden.HERE / f"{name}_golden.npz", allow_pickle=False) as g:
        golden = {k: g[k] for k in g.files}
    packed, cfg, iterations = make_golden.case_inputs(name)
    assert np.array_equal(packed.placed, golden["placed"])
    with _native.World(packed, cfg) as w:
        w.rebuild_lists()
        counts, offsets, segments = w.cell_list()
        assert np.array_equal(counts, golden["cell_counts"])
        assert np.array_equal(offsets, golden["cell_offsets"])
        assert np.array_equal(segments, golden["cell_segments"])
        ncount, lists = w.neighbor_lists()
        assert np.array_equal(ncount, golden["neighbor_counts"])
        block = lists.shape[1]
        for s in np.nonzero(ncount <= block)[0]:
            assert np.array_equal(lists[s, :ncount[s]], golden["neighbor_lists"][s, :ncount[s]]), s
        status = w.run_batch(cfg, iterations)
        assert status.total_iterations == int(golden["total_iterations"][0])
        positions = w.download_positions()
    d = 10e-6 if name == "config1" else 7e-6
    assert np.max(np.abs(positions.astype(np.float64) - golden["positions"])) <= 1e-4 * d
    assert np.array_equal(positions.view(np.uint32), golden["positions"].view(np.uint32))
