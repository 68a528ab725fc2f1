"""tgh_mlp_forward_static, the forward pass that allocates no device memory
(its staging is a static device array of the library), against
tgh_mlp_forward through the C ABI: bit for bit on two calls in a row, the
second with fewer rows, and a clear error beyond its row limit."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import mlp_ref
from tangle_amd import _native, learned

pytestmark = pytest.mark.gpu


def _features(rows, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2.0, 2.0, (rows, 64)).astype(np.float32)
    x[:, 51:] = 0.0
    return np.ascontiguousarray(mlp_ref.f32_to_bf16_bits(x))


def _forward(lib, fn, model, x):
    rows = x.shape[0]
    out = np.full((rows, 3), np.nan, dtype=np.float32)
    _native.check(getattr(lib, fn)(model._handle, _native.ptr(x, C.c_uint16), rows, _native.ptr(out, C.c_float)))
    return out


def test_static_forward_equals_forward(hip_device):
    lib = hip_device
    model = learned.WarmStartModel(seed=11)
    try:
        for rows, seed in ((50_000, 1), (1_237, 2)):
            print(f"[case] forward_static rows={rows}", flush=True)
            x = _features(rows, seed)
            a = _forward(lib, "tgh_mlp_forward", model, x)
            b = _forward(lib, "tgh_mlp_forward_static", model, x)
            assert np.isfinite(b).all()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rows
        limit = _native.TGH_MLP_FORWARD_STATIC_MAX_ROWS
        x = _features(limit, 3)
        assert np.array_equal(_forward(lib, "tgh_mlp_forward", model, x).view(np.uint32),
                              _forward(lib, "tgh_mlp_forward_static", model, x).view(np.uint32))
        with pytest.raises(ValueError, match="at most"):
            _forward(lib, "tgh_mlp_forward_static", model, _features(limit + 1, 4))
    finally:
        model.close()
