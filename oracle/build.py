"""Build recipe for the CPU oracle (liboracle.so).

The reference itself (CUDA C++, nvcc-only, kernels launched on the legacy
default stream) cannot be built in this image, so there is no oracle/_ref;
the oracle is this restatement alone.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    out = os.path.join(HERE, "liboracle.so")
    # x86-64-v3 (AVX2+FMA) so the library built here also runs on the
    # benchmark host; fp contraction off: every fma is explicit.
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
           "-std=gnu11", "-fPIC", "-shared", "-pthread",
           os.path.join(HERE, "oracle.c"), "-o", out + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
