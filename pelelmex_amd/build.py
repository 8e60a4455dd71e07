"""Build the HIP library in-tree: pelelmex_amd/libpelelmex_amd.so (gfx950)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpelelmex_amd.so")
ARCH = os.environ.get("PR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of generated code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE expression trees as written: results match the CPU oracle bit for bit
    "-ffp-contract=off",
]


def build_library(verbose=False):
    src = os.path.join(CSRC, "reactor.hip")
    tmp = LIB + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
