"""The oracle against the committed golden vectors (regression pins written
by tests/golden/make_golden.py): generated inputs, cell list, neighbour lists
and the state after a fixed number of iterations, all bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from golden import make_golden


@pytest.mark.parametrize("name", list(make_golden.CASES))
def test_oracle_reproduces_golden(name):
    with np.load(make_golden.HERE / f"{name}_golden.npz", allow_pickle=False) as g:
        golden = {k: g[k] for k in g.files}
    got = make_golden.compute(name)
    assert set(got) == set(golden)
    for key in golden:
        a, b = np.asarray(got[key]), golden[key]
        assert a.shape == b.shape, key
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), key
