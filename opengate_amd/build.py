"""Compile the HIP library (libgatehip.so) for gfx950, in-tree.

Usage: python -m opengate_amd.build  (or call build()).
The library is placed next to this file so that it travels with the source tree.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_NAME = "libgatehip.so"
LIB_PATH = PKG / LIB_NAME

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GH_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the transport relies on an explicit, reproducible FP32 operation
# order (every fused multiply-add is written as fmaf); packed FP32 instructions are
# disabled on purpose.
COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose=False):
    srcs = [str(s) for s in sources()]
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [HIPCC, *COMMON_FLAGS, "-shared", "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB_PATH)
