"""The curvature sweep's arc cosine (tangle_amd/csrc/acosf_f32.h) against the
host libm's acosf, bit for bit (CPU). The header is compiled with g++ into a
small harness (the comparison was also run over every binary32 value in
[-1, 1], 2,130,706,434 of them, with no difference); it is checked on every
61st binary32 pattern in [-1, 1], on
the neighbourhoods of the case boundaries (0, +-1/2, +-1, 2^-26) and on the
closed forms acos(0) = pi/2, acos(1) = 0, acos(-1) = pi, acos(1/2) = pi/3."""

from __future__ import annotations

import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

HEADER = Path(__file__).resolve().parents[1] / "tangle_amd" / "csrc" / "acosf_f32.h"

HARNESS = r'''
#include <stdint.h>
#include <string.h>
#include "acosf_f32.h"
extern "C" void eval(const float *x, float *out, long n) {
    for (long i = 0; i < n; ++i) out[i] = tgh::acosf_f32(x[i]);
}
extern "C" void libm(const float *x, float *out, long n) {
    for (long i = 0; i < n; ++i) out[i] = acosf(x[i]);
}
'''


@pytest.fixture(scope="module")
def acos_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("acosf")
    (d / "h.cpp").write_text(HARNESS)
    so = d / "libh.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-builtin", "-shared", "-fPIC", f"-I{HEADER.parent}",
                    str(d / "h.cpp"), "-o", str(so), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    lib.eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.libm.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return lib


def evaluate(lib, x, fn="eval"):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    getattr(lib, fn)(x.ctypes.data, out.ctypes.data, x.size)
    return out


def libm_acosf(x):
    libm = ctypes.CDLL("libm.so.6")
    libm.acosf.restype = ctypes.c_float
    libm.acosf.argtypes = [ctypes.c_float]
    return np.array([libm.acosf(float(v)) for v in x], dtype=np.float32)


def test_closed_forms(acos_lib):
    got = evaluate(acos_lib, [0.0, 1.0, -1.0, 0.5, -0.5])
    assert got[0] == np.float32(math.pi / 2) and got[1] == 0.0
    assert abs(got[2] - math.pi) < 3e-7 and abs(got[3] - math.pi / 3) < 1.2e-7
    assert abs(got[4] - 2 * math.pi / 3) < 2.4e-7


def test_matches_libm_on_boundaries(acos_lib):
    centres = np.float32([0.0, 0.5, -0.5, 1.0, -1.0, 2.0 ** -26, -(2.0 ** -26)])
    u = centres.view(np.uint32).astype(np.int64)
    offsets = np.arange(-2000, 2001)
    pats = (u[:, None] + offsets[None, :]).ravel()
    pats = pats[(pats >= 0) & (pats < 2 ** 32)].astype(np.uint32)
    x = pats.view(np.float32)
    x = x[np.abs(x) <= 1.0]
    assert np.array_equal(evaluate(acos_lib, x).view(np.uint32), libm_acosf(x).view(np.uint32))


def test_matches_libm_on_a_stride(acos_lib):
    # positive and negative halves of [-1, 1], every 61st bit pattern
    pos = np.arange(0, 0x3F800001, 61, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([pos.view(np.float32), (pos | np.uint32(0x80000000)).view(np.float32)])
    got = evaluate(acos_lib, x)
    ref = evaluate(acos_lib, x, "libm")
    assert np.array_equal(ref[:: 4099].view(np.uint32), libm_acosf(x[:: 4099]).view(np.uint32))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
