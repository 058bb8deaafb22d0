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

ENTRY_POINTS = [  # C++ symbols the reference's callers link against

    "Batched_ILU_Preconditioned_BiCGSTAB_Gpu", "Batched_ILU_ISAI_Preconditioned_BiCGSTAB_Gpu",
    "Batched_General_ISAI_Preconditioned_BiCGSTAB_Gpu", "Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu",
    "Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_segregated_Gpu", "ILU_0_Factorization_Gpu", "ComputeILU0Approach1",
    "ParILU_0_Factorization_Gpu", "GenerateISAI_gpu", "PagedCSRMatrices::CopyFromCpuToGpu",
    "PagedVectors::CopyFromGpuToCpu", "SolverResults::SolverResults", "Batched_Preconditioned_BiCGSTAB_Gpu",
    "FillSubdir", "ReadData", "Print_ans", "PrintSolverResults",
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


def ilu_isai_power(pat):
    """the drop-in ILU-ISAI solver's power: the reference's 2 when |L|^2, |U|^2
    rows fit 32 entries, else 1"""
    import oracle as O
    for typ in (O.ISAI_LOWER, O.ISAI_UPPER):
        mrp, _ = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 2)
        if np.diff(mrp).max() > 32:
            return 1
    return 2


def dropin_oracle(pat, A, b, variant):
    """the oracle's results for a drop-in solver variant (ATOL 1e-11 absolute, MAX_ITER 500)"""
    import oracle as O
    from test_gpu_parity import run_oracle
    precond, kernel = VARIANTS[variant]
    if variant in ("ilu_isai", "parilu_isai") and ilu_isai_power(pat) == 2:
        opc = O.ILU0_ISAI if variant == "ilu_isai" else O.PARILU0_ISAI
        return O.solve_isai_pow(pat, A, b, opc, 5, 2, O.opts(tol=1e-11, tol_mode=0))
    return run_oracle(pat, A, b, precond=precond, parilu_sweeps=5, tol=1e-11, tol_mode=0, kernel=kernel)


VARIANTS = {"merged": (bilu.ILU0, bilu.KERNEL_AUTO), "segregated": (bilu.ILU0, bilu.KERNEL_AUTO),
            "ilu": (bilu.ILU0, bilu.KERNEL_AUTO), "parilu": (bilu.PARILU0, bilu.KERNEL_AUTO),
            "ilu_isai": (bilu.ILU0_ISAI, bilu.KERNEL_AUTO), "parilu_isai": (bilu.PARILU0_ISAI, bilu.KERNEL_AUTO),
            "general_isai": (bilu.GENERAL_ISAI, bilu.KERNEL_AUTO)}


@pytest.mark.gpu
@pytest.mark.parametrize("kind,B,n", [("stencil5", 64, 100), ("random", 128, 144)])
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_driver_matches_oracle(tmp_path, kind, B, n, variant):
    print(f"case drop-in {variant} {kind} B={B} n={n}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, pat, A, b)
    r = subprocess.run([DRIVER, variant, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x, it, cv, ir, tr, t = read_result(fout, B, pat.n)
    # the reference's semantics: absolute ATOL = 1e-11, MAX_ITER = 500, no restart; the
    # ILU-ISAI solvers on |L|^2, |U|^2 where those rows fit 32 entries (stencil), else power 1
    o = dropin_oracle(pat, A, b, variant)
    for name, u, v in zip(["x", "iters", "conv", "iter_res", "true_res"], (x, it, cv, ir, tr), o):
        assert np.array_equal(u, v), f"{variant}: {name} differs"
    PST, IST, OET = t
    assert OET > 0 and IST > 0 and OET >= IST
    if variant in ("ilu", "parilu"):
        assert PST > 0 and OET == pytest.approx(PST + IST)


def read_factors(path, B, n):
    raw = open(path, "rb").read()
    o = 0

    def take(dt, cnt):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=cnt, offset=o)
        o += a.nbytes
        return a
    ln, un = take(np.int32, 2)
    lrp, lci = take(np.int32, n + 1), take(np.int32, ln)
    urp, uci = take(np.int32, n + 1), take(np.int32, un)
    return lrp, lci, urp, uci, take(np.float64, B * ln).reshape(B, ln), take(np.float64, B * un).reshape(B, un)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["ilu0", "parilu0", "ilu0_app2", "ilu0_app3", "ilu0_app4"])
def test_driver_factorizations_match_oracle(tmp_path, variant):
    import oracle as O
    print(f"case drop-in factor {variant}", flush=True)
    pat, A, b = synth.batch("random", 64, n=144)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, pat, A, b)
    r = subprocess.run([DRIVER, variant, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lrp, lci, urp, uci, Lv, Uv = read_factors(fout, 64, pat.n)
    rp, ci, diag = pat.row_ptrs, pat.col_idx, pat.diag
    for s in range(64):
        F = O.ilu0(pat, A[s])[0] if variant.startswith("ilu0") else O.parilu0(pat, A[s], 5)
        # reference split: L = strict lower + explicit 1 on the diagonal, U = diagonal + upper
        Lw, Uw, lc, uc = [], [], [], []
        for i in range(pat.n):
            for p in range(rp[i], diag[i]):
                Lw.append(F[p]); lc.append(ci[p])
            Lw.append(1.0); lc.append(i)
            for p in range(diag[i], rp[i + 1]):
                Uw.append(F[p]); uc.append(ci[p])
        assert np.array_equal(lci, lc) and np.array_equal(uci, uc)
        assert np.array_equal(Lv[s], Lw) and np.array_equal(Uv[s], Uw)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,typ", [("isai_general", 0), ("isai_lower", 1), ("isai_upper", 2)])
def test_driver_isai_matches_oracle(tmp_path, variant, typ):
    import oracle as O
    print(f"case drop-in {variant}", flush=True)
    pat, A, b = synth.batch("random", 32, n=90)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, pat, A, b)
    r = subprocess.run([DRIVER, variant, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    M = np.fromfile(fout, dtype=np.float64).reshape(32, pat.nnz)
    rp, diag = pat.row_ptrs, pat.diag
    for s in range(32):
        want, st = O.isai(pat, A[s], typ)
        assert st == 0
        if typ == 0:
            assert np.array_equal(M[s], want)
        else:
            mask = np.zeros(pat.nnz, bool)
            for i in range(pat.n):
                rng = range(rp[i], diag[i] + 1) if typ == 1 else range(diag[i], rp[i + 1])
                mask[list(rng)] = True
            assert np.array_equal(M[s][mask], want[mask])
            assert np.all(M[s][~mask] == 0.0)


@pytest.mark.gpu
def test_driver_isai_lower_nonunit_diagonal(tmp_path):
    """GenerateISAI_gpu(lower_tri) of a lower-triangular matrix whose stored
    diagonal is not 1 divides by it, as the reference's lower solve does
    (isai_gpu.cu:359-389): the oracle's ISAI of the matrix on its own pattern,
    bit for bit, and (M L)(i, j) = delta_ij on the pattern"""
    import oracle as O
    print("case drop-in isai lower, non-unit diagonal", flush=True)
    full, A, b = synth.batch("random", 16, n=90)
    keep = full.col_idx <= full.rows
    rp = np.concatenate([[0], np.cumsum(np.bincount(full.rows[keep], minlength=full.n))]).astype(np.int32)
    lpat = synth.Pattern(full.n, rp, full.col_idx[keep])
    L = np.ascontiguousarray(A[:, keep])
    L[:, lpat.diag] *= np.linspace(0.5, 3.0, L.shape[0])[:, None]     # stored diagonal far from 1
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, lpat, L, b)
    r = subprocess.run([DRIVER, "isai_lower", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    M = np.fromfile(fout, dtype=np.float64).reshape(16, lpat.nnz)
    for s in range(16):
        want, st = O.isai(lpat, L[s], 0)
        assert st == 0 and np.array_equal(M[s], want)
        P = lpat.dense(M[s]) @ lpat.dense(L[s])
        I = np.eye(lpat.n)
        assert np.abs((P - I)[lpat.rows, lpat.col_idx]).max() < 1e-10
