"""Builds the HIP relaxation library (libtangle_hip.so) for gfx950 in-tree.

The library is compiled with hipcc as one shared object from the kernel
translation unit and the host sources. Float arithmetic is compiled without
contraction and with IEEE division/square root so the kernels reproduce the
reference's f32 results (cell lists and neighbour pair sets bit for bit).

Usage: python -m tangle_amd.build   (or tangle_amd.build.build_library())
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIB_DIR = PACKAGE / "_lib"
LIBRARY = LIB_DIR / "libtangle_hip.so"

# (source, language): host code that calls the HIP runtime is compiled as HIP
SOURCES = [
    ("relax_kernels.hip", "hip"),
    ("learned.hip", "hip"),
    ("world.cpp", "hip"),
    ("learned_host.cpp", "hip"),
    ("population.cpp", "c++"),
]
ARCH = os.environ.get("TANGLE_HIP_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 instructions are kept out of the device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build tangle_amd")
    return path


def build_library(verbose: bool = False) -> Path:
    """Compiles every source and links libtangle_hip.so; returns its path."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    build_dir = LIB_DIR / "obj"
    build_dir.mkdir(parents=True, exist_ok=True)
    objects = []
    for name, language in SOURCES:
        src = CSRC / name
        obj = build_dir / (name + ".o")
        if language == "hip":
            cmd = [hipcc(), *COMMON_FLAGS, *DEVICE_FLAGS, "-c", "-x", "hip", str(src), "-o", str(obj)]
        else:
            cmd = [hipcc(), *COMMON_FLAGS, "-c", "-x", "c++", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objects.append(str(obj))
    tmp = LIBRARY.with_suffix(".so.tmp")
    link = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objects]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    build_library(verbose="-v" in sys.argv)
    print(LIBRARY)
