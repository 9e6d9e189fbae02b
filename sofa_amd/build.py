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
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# translation units: (extra flags, object name).  The FP64 engines and the C ABI keep every
# multiply and add separate (the reference's rounding); the FP32 engine may fuse them.
UNITS = [
    (["-ffp-contract=off"], "engine_fp64.o"),
    (["-ffp-contract=fast", "-DSOFA_AMD_TU_FP32"], "engine_fp32.o"),
]


def build(verbose: bool = False) -> str:
    import concurrent.futures as cf
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmds = []
    for extra, obj in UNITS:
        cmds.append([HIPCC, *FLAGS, *extra, "-I", os.path.join(ROOT, "include"), "-c", *SOURCES,
                     "-o", os.path.join(bdir, obj)])
    with cf.ThreadPoolExecutor(len(cmds)) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        list(ex.map(subprocess.check_call, cmds))
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[os.path.join(bdir, o) for _, o in UNITS],
            "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.check_call(link)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
