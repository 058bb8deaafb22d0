"""Build recipe for libbilu.so (HIP, gfx950) -- compiled in-tree so the
shared library travels with the source tree."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbilu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every fused multiply-add in the kernels is an explicit
# fma(), so results are reproducible bit for bit by the CPU restatement.
# packed FP32 ops stay disabled (the compiler may not form v_pk_* FP32).
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def build(verbose: bool = False, extra=()) -> str:
    srcs = [os.path.join(CSRC, "solver.hip"), os.path.join(CSRC, "mlp.hip"), os.path.join(CSRC, "pattern.cpp")]
    cmd = [HIPCC, *HIP_FLAGS, *extra, "-shared", *srcs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
