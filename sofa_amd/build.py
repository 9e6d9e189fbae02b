"""Builds libsofa_amd.so in-tree with hipcc for gfx950.

Flags:
* ``--offload-arch=gfx950`` only (MI355X).
* ``-ffp-contract=off``: no fused multiply-add, so the FP64 kernels reproduce the reference's
  operation-by-operation rounding.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsofa_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "engine.hip")]
HEADERS = [os.path.join(HERE, "csrc", h) for h in ("fem_device.h", "fem_kernels.h")] + \
    [os.path.join(ROOT, "include", "sofa_amd.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), *SOURCES, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
