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


CPP = os.path.join(HERE, "cpp")
LIB_CPP = os.path.join(HERE, "libbilu_cpp.so")
DRIVER = os.path.join(HERE, "bicgstab_driver")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, extra=()) -> str:
    srcs = [os.path.join(CSRC, "solver.hip"), os.path.join(CSRC, "mlp.hip"), os.path.join(CSRC, "pattern.cpp")]
    _run([HIPCC, *HIP_FLAGS, *extra, "-shared", *srcs, "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    # drop-in C++ interface (host code over the C ABI) and its driver program
    host = ["-O2", "-std=c++17", "-fPIC", "-Wno-unused-command-line-argument"]
    link = ["-L" + HERE, "-lbilu", "-Wl,-rpath,$ORIGIN"]
    _run([HIPCC, *host, "-shared", os.path.join(CPP, "containers.cpp"), os.path.join(CPP, "drop_in.cpp"),
          os.path.join(CPP, "ReadWriteData.cpp"), *link, "-o", LIB_CPP + ".tmp"], verbose)
    os.replace(LIB_CPP + ".tmp", LIB_CPP)
    _run([HIPCC, *host, os.path.join(CPP, "bicgstab_driver.cpp"), "-L" + HERE, "-lbilu_cpp", "-lbilu",
          "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"], verbose)
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
