"""Writes the golden fixtures in this folder from the CPU oracle.

The reference ships no test data, so these vectors are the oracle's own
outputs on the synthetic workloads, committed so that oracle regressions and
GPU results can be checked without recomputing.  Inputs are regenerated from
the counter-based generator (synth.py), so only outputs are stored.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "merging-ilu0_amd")]

import oracle as O  # noqa: E402
import synth  # noqa: E402

CASES = {
    # name: (kind, batch, n)
    "stencil5_b64": ("stencil5", 64, 100),
    "random_n144_b8": ("random", 8, 144),
    "random_n32_b8": ("random", 8, 32),
    "random_n300_b4": ("random", 4, 300),
}


def make(name, kind, B, n):
    pat, A, b = synth.batch(kind, B, n=n)
    out = {"row_ptrs": pat.row_ptrs, "col_idx": pat.col_idx}
    for tag, o in (("ref", O.reference_opts()), ("w64", O.opts()),
                   ("rel", O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1))):
        x, it, cv, ir, tr = O.batch_solve(pat, A, b, O.ILU0, o=o)
        out[f"x_{tag}"], out[f"iter_{tag}"], out[f"conv_{tag}"] = x, it, cv
        out[f"iter_res_{tag}"], out[f"true_res_{tag}"] = ir, tr
    k = min(B, 4)
    out["ilu0"] = np.stack([O.ilu0(pat, A[s])[0] for s in range(k)])
    out["parilu5"] = np.stack([O.parilu0(pat, A[s], 5) for s in range(k)])
    F0 = out["ilu0"][0]
    out["isai_lower"] = O.isai(pat, F0, O.ISAI_LOWER)[0]
    out["isai_upper"] = O.isai(pat, F0, O.ISAI_UPPER)[0]
    if max(np.diff(pat.row_ptrs)) <= 32:
        out["isai_general"] = O.isai(pat, A[0], O.ISAI_GENERAL)[0]
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, pat.nnz, "iters(ref)", out["iter_ref"].min(), out["iter_ref"].max())


if __name__ == "__main__":
    for name, (kind, B, n) in CASES.items():
        make(name, kind, B, n)
