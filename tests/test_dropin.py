"""Drop-in C++ interface (libbilu_cpp.so + the bicgstab_driver program):
the reference-named entry points exist (CPU), and through the driver they
give the oracle's results bit for bit (GPU)."""
import os
import subprocess

import numpy as np
import pytest

import bilu
import synth

HERE = os.path.dirname(os.path.abspath(bilu.__file__))
LIB_CPP = os.path.join(HERE, "libbilu_cpp.so")
DRIVER = os.path.join(HERE, "bicgstab_driver")

ENTRY_POINTS = [
    "Batched_ILU_Preconditioned_BiCGSTAB_Gpu", "Batched_ILU_ISAI_Preconditioned_BiCGSTAB_Gpu",
    "Batched_General_ISAI_Preconditioned_BiCGSTAB_Gpu", "Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu",
    "Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_segregated_Gpu", "ILU_0_Factorization_Gpu", "ComputeILU0Approach1",
    "ParILU_0_Factorization_Gpu", "GenerateISAI_gpu", "PagedCSRMatrices::CopyFromCpuToGpu",
    "PagedVectors::CopyFromGpuToCpu", "SolverResults::SolverResults", "Batched_Preconditioned_BiCGSTAB_Gpu",
]


def test_cpp_library_exports_reference_entry_points():
    if not os.path.exists(LIB_CPP):
        pytest.skip("libbilu_cpp.so not built")
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB_CPP], capture_output=True, text=True, check=True).stdout
    for name in ENTRY_POINTS:
        assert name + "(" in out, name


def write_problem(path, pat, A, b):
    with open(path, "wb") as f:
        np.array([pat.n, pat.nnz, A.shape[0]], np.int32).tofile(f)
        pat.row_ptrs.astype(np.int32).tofile(f)
        pat.col_idx.astype(np.int32).tofile(f)
        np.ascontiguousarray(A, np.float64).tofile(f)
        np.ascontiguousarray(b, np.float64).tofile(f)


def read_result(path, B, n):
    raw = open(path, "rb").read()
    o = 0

    def take(dt, cnt):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=cnt, offset=o)
        o += a.nbytes
        return a
    x = take(np.float64, B * n).reshape(B, n)
    it, cv = take(np.float32, B), take(np.int32, B)
    ir, tr = take(np.float64, B), take(np.float64, B)
    t = take(np.float64, 3)
    return x, it, cv, ir, tr, t


VARIANTS = {"merged": (bilu.ILU0, bilu.KERNEL_AUTO), "segregated": (bilu.ILU0, bilu.KERNEL_AUTO),
            "ilu": (bilu.ILU0, bilu.KERNEL_AUTO), "parilu": (bilu.PARILU0, bilu.KERNEL_AUTO),
            "ilu_isai": (bilu.ILU0_ISAI, bilu.KERNEL_AUTO), "parilu_isai": (bilu.PARILU0_ISAI, bilu.KERNEL_AUTO),
            "general_isai": (bilu.GENERAL_ISAI, bilu.KERNEL_AUTO)}


@pytest.mark.gpu
@pytest.mark.parametrize("kind,B,n", [("stencil5", 64, 100), ("random", 128, 144)])
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_driver_matches_oracle(tmp_path, kind, B, n, variant):
    from test_gpu_parity import run_oracle
    print(f"case drop-in {variant} {kind} B={B} n={n}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, pat, A, b)
    r = subprocess.run([DRIVER, variant, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x, it, cv, ir, tr, t = read_result(fout, B, pat.n)
    precond, kernel = VARIANTS[variant]
    # the reference's semantics: absolute ATOL = 1e-11, MAX_ITER = 500, no restart
    o = run_oracle(pat, A, b, precond=precond, parilu_sweeps=5, tol=1e-11, tol_mode=0, kernel=kernel)
    for name, u, v in zip(["x", "iters", "conv", "iter_res", "true_res"], (x, it, cv, ir, tr), o):
        assert np.array_equal(u, v), f"{variant}: {name} differs"
    PST, IST, OET = t
    assert OET > 0 and IST > 0 and OET >= IST
    if variant in ("ilu", "parilu"):
        assert PST > 0 and OET == pytest.approx(PST + IST)
