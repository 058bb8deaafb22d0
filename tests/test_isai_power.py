"""ISAI with power-k sparsity patterns (reference GenerateISAI_gpu(aiA, A,
type, power), isai_gpu.cu:40-178 pattern, :559-626 values; the reference's
ILU-ISAI solver uses power 2, PrecondBiCGSTAB_ilu_isai_gpu.cu:297-299).

CPU: the oracle's pattern is pinned by dense boolean matrix powers; its
values by the defining identity (M T)(i, J_i) = e_i, by the exact inverse
when the power pattern closes the triangle, and by the power-1 route equal
to the pinned power-1 ISAI bit for bit (generation and whole solves).
GPU: the library's patterns, values and solves equal the oracle's bit for
bit; the drop-in's GenerateISAI_gpu(power 2) and ILU-ISAI solver follow."""
import os
import subprocess

import numpy as np
import pytest

import bilu
import oracle as O
import synth


def typed_dense_pattern(pat, typ):
    D = pat.dense(np.ones(pat.nnz)) != 0
    if typ == O.ISAI_LOWER:
        D = np.tril(D, -1) | np.eye(pat.n, dtype=bool)
    elif typ == O.ISAI_UPPER:
        D = np.triu(D)
    return D


def typed_dense_values(pat, vals, typ):
    D = pat.dense(vals)
    if typ == O.ISAI_LOWER:
        D = np.tril(D, -1) + np.eye(pat.n)
    elif typ == O.ISAI_UPPER:
        D = np.triu(D)
    return D


def pattern_cases():
    yield synth.batch("stencil5", 1, n=100)[0]
    yield synth.batch("random", 1, n=32)[0]
    yield synth.batch("random", 1, n=90)[0]


@pytest.mark.parametrize("typ", [O.ISAI_GENERAL, O.ISAI_LOWER, O.ISAI_UPPER])
@pytest.mark.parametrize("power", [1, 2, 3])
def test_oracle_pattern_is_the_boolean_power(typ, power):
    for pat in pattern_cases():
        P = typed_dense_pattern(pat, typ).astype(np.int64)
        Q = P.copy()
        for _ in range(power - 1):
            Q = ((Q @ P) > 0).astype(np.int64)
        mrp, mci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, power)
        want = [np.nonzero(Q[i])[0] for i in range(pat.n)]
        assert np.array_equal(np.diff(mrp), [len(w) for w in want])
        assert np.array_equal(mci, np.concatenate(want))


@pytest.mark.parametrize("typ", [O.ISAI_GENERAL, O.ISAI_LOWER, O.ISAI_UPPER])
@pytest.mark.parametrize("kind,n,power", [("stencil5", 100, 2), ("stencil5", 100, 3), ("random", 32, 2),
                                          ("random", 144, 2)])
def test_oracle_values_satisfy_the_isai_identity(typ, kind, n, power):
    pat, A, _ = synth.batch(kind, 3, n=n)
    mrp, mci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, power)
    if typ == O.ISAI_GENERAL and np.diff(mrp).max() > 32:
        pytest.skip("general ISAI rows beyond 32 entries")
    for s in range(3):
        vals = A[s] if typ == O.ISAI_GENERAL else O.ilu0(pat, A[s])[0]
        M, st = O.isai_on(pat, vals, typ, mrp, mci)
        assert st == 0
        T = typed_dense_values(pat, vals, typ)
        MT = synth.Pattern(pat.n, mrp, mci).dense(M) @ T
        for i in range(pat.n):
            J = mci[mrp[i]:mrp[i + 1]]
            e = (J == i).astype(float)
            assert np.allclose(MT[i, J], e, rtol=0, atol=1e-12 * np.abs(T).max() * np.abs(M).max()), (s, i)


@pytest.mark.parametrize("typ", [O.ISAI_LOWER, O.ISAI_UPPER])
def test_oracle_closed_power_pattern_gives_the_exact_inverse(typ):
    # 4x4 grid: the 6th power reaches the transitive closure of the triangle's
    # graph, which holds the support of its inverse; M is then the inverse
    pat = synth.stencil5_pattern(4, 4)
    A = synth.stencil5_values(pat, np.arange(2))
    mrp, mci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 6)
    m5, _ = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 5)
    m7, c7 = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 7)
    assert m5[-1] < mrp[-1] and np.array_equal(m7, mrp) and np.array_equal(c7, mci)
    support = synth.Pattern(pat.n, mrp, mci).dense(np.ones(len(mci))) != 0
    for s in range(2):
        F = O.ilu0(pat, A[s])[0]
        M, st = O.isai_on(pat, F, typ, mrp, mci)
        assert st == 0
        inv = np.linalg.inv(typed_dense_values(pat, F, typ))
        assert np.all(np.abs(inv[~support]) <= 1e-15 * np.abs(inv).max())
        Md = synth.Pattern(pat.n, mrp, mci).dense(M)
        assert np.allclose(Md, inv, rtol=1e-12, atol=1e-14 * np.abs(inv).max())


@pytest.mark.parametrize("typ", [O.ISAI_GENERAL, O.ISAI_LOWER, O.ISAI_UPPER])
def test_oracle_power_one_route_equals_the_pinned_isai(typ):
    pat, A, _ = synth.batch("random", 4, n=90)
    mrp, mci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 1)
    for s in range(4):
        vals = A[s] if typ == O.ISAI_GENERAL else O.ilu0(pat, A[s])[0]
        want, st0 = O.isai(pat, vals, typ)
        got, st1 = O.isai_on(pat, vals, typ, mrp, mci)
        assert st0 == st1 == 0
        # power-1 rows are windows of the pattern: compare through the typed positions
        rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
        keep = {O.ISAI_GENERAL: np.ones(pat.nnz, bool), O.ISAI_LOWER: pat.col_idx <= rows,
                O.ISAI_UPPER: pat.col_idx >= rows}[typ]
        assert np.array_equal(got, want[keep])


@pytest.mark.parametrize("precond", [O.ILU0_ISAI, O.PARILU0_ISAI, O.GENERAL_ISAI])
def test_oracle_power_one_solve_equals_the_batch_solver(precond):
    pat, A, b = synth.batch("stencil5", 8, n=100)
    want = O.batch_solve(pat, A, b, precond=precond)
    got = O.solve_isai_pow(pat, A, b, precond, 5, 1)
    for u, v in zip(got, want):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("power", [2, 3])
def test_oracle_power_k_solves_converge(power):
    pat, A, b = synth.batch("stencil5", 8, n=100)
    x, it, cv, ir, tr = O.solve_isai_pow(pat, A, b, O.ILU0_ISAI, 5, power,
                                         O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1))
    assert np.all(cv == 1)
    for s in range(8):   # independent residual check
        r = b[s] - pat.dense(A[s]) @ x[s]
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b[s]) * 1.0001
    it1 = O.solve_isai_pow(pat, A, b, O.ILU0_ISAI, 5, 1)[1]
    assert it.mean() <= it1.mean()   # the wider pattern never needs more iterations here


# ------------------------------------------------------------------ GPU

GPU_CASES = [("stencil5", 64, 100, 2), ("stencil5", 32, 100, 3), ("random", 64, 32, 2)]
# ISAI(L) / ISAI(U) rows of 84 / 86 entries (one row per wave, two unknowns per lane)
LONG_CASES = [("random", 16, 144, 2)]


def row_limit(typ):
    """longest ISAI row the library takes: 32 for the general ISAI, 128 for the triangles"""
    return 32 if typ == bilu.ISAI_GENERAL else 128


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [bilu.ISAI_GENERAL, bilu.ISAI_LOWER, bilu.ISAI_UPPER])
@pytest.mark.parametrize("kind,B,n,power", GPU_CASES + LONG_CASES)
def test_library_isai_power_matches_oracle(kind, B, n, power, typ):
    import torch
    print(f"case isai power {power} type {typ} {kind} B={B} n={n}", flush=True)
    pat, A, _ = synth.batch(kind, B, n=n)
    P = bilu.Pattern.from_synth(pat)
    mrp, mci = bilu.isai_pattern(P, typ, power)
    wrp, wci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, power)
    assert np.array_equal(mrp, wrp) and np.array_equal(mci, wci)
    vals = A if typ == bilu.ISAI_GENERAL else np.stack([O.ilu0(pat, A[s])[0] for s in range(B)])
    dev = torch.device("cuda:0")
    dv = torch.from_numpy(vals).to(dev)
    out = torch.full((B, len(mci)), float("nan"), dtype=torch.float64, device=dev)
    st = torch.empty(B, dtype=torch.int32, device=dev)
    if np.diff(mrp).max() > row_limit(typ):
        with pytest.raises(bilu.BiluError):
            bilu.isai_power(P, dv, typ, power, out, st)
        return
    bilu.isai_power(P, dv, typ, power, out, st)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(st.cpu().numpy() == 0)
    for s in range(B):
        want, ost = O.isai_on(pat, vals[s], typ, mrp, mci)
        assert ost == 0 and np.array_equal(got[s], want), s


@pytest.mark.gpu
def test_library_isai_power_rejects_long_rows():
    import torch
    # ISAI(L) rows beyond 128 entries (|L|^3 at n = 300: 193) and general rows beyond 32 (|A|^2 at n = 144)
    for n, typ, power, pc in ((300, bilu.ISAI_LOWER, 3, bilu.ILU0_ISAI), (144, bilu.ISAI_GENERAL, 2, bilu.GENERAL_ISAI)):
        pat, A, _ = synth.batch("random", 4, n=n)
        P = bilu.Pattern.from_synth(pat)
        mrp, _ = bilu.isai_pattern(P, typ, power)
        assert np.diff(mrp).max() > row_limit(typ)
        dv = torch.from_numpy(A).cuda()
        out = torch.empty((4, int(mrp[-1])), dtype=torch.float64, device="cuda")
        with pytest.raises(bilu.BiluError, match="4"):
            bilu.isai_power(P, dv, typ, power, out)
        o = bilu.opts(precond=pc, isai_power=power)
        with pytest.raises(bilu.BiluError):
            bilu.solve(P, dv, torch.ones(4, pat.n, dtype=torch.float64, device="cuda"),
                       torch.empty(4, pat.n, dtype=torch.float64, device="cuda"), o)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", [bilu.ILU0_ISAI, bilu.PARILU0_ISAI, bilu.GENERAL_ISAI])
@pytest.mark.parametrize("kind,B,n,power", GPU_CASES + LONG_CASES)
def test_solve_with_isai_power_matches_oracle(kind, B, n, power, precond):
    import torch
    print(f"case solve isai power {power} precond {precond} {kind} B={B} n={n}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    if precond == bilu.GENERAL_ISAI:
        mrp, _ = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, O.ISAI_GENERAL, power)
        if np.diff(mrp).max() > 32:
            pytest.skip("general ISAI rows beyond 32 entries")
    P = bilu.Pattern.from_synth(pat)
    dev = torch.device("cuda:0")
    dA, db = torch.from_numpy(A).to(dev), torch.from_numpy(b).to(dev)
    x = torch.empty_like(db)
    o = bilu.opts(precond=precond, isai_power=power, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, true_res_restart=1)
    it, cv, ir, tr = bilu.solve(P, dA, db, x, o)
    torch.cuda.synchronize()
    g = (x.cpu().numpy(), it.cpu().numpy(), cv.cpu().numpy(), ir.cpu().numpy(), tr.cpu().numpy())
    opc = {bilu.ILU0_ISAI: O.ILU0_ISAI, bilu.PARILU0_ISAI: O.PARILU0_ISAI, bilu.GENERAL_ISAI: O.GENERAL_ISAI}[precond]
    w = O.solve_isai_pow(pat, A, b, opc, 5, power, O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1))
    for name, u, v in zip(["x", "iters", "conv", "iter_res", "true_res"], g, w):
        assert np.array_equal(u, v), name
    # power 1 through the same option keeps the power-1 kernels' results
    if precond != bilu.GENERAL_ISAI and kind == "stencil5" and power == 2:
        o1 = bilu.opts(precond=precond, isai_power=1, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, true_res_restart=1)
        it1 = bilu.solve(P, dA, db, x, o1)[0]
        assert it.cpu().numpy().mean() <= it1.cpu().numpy().mean()


HERE = os.path.dirname(os.path.abspath(bilu.__file__))
DRIVER = os.path.join(HERE, "bicgstab_driver")


@pytest.mark.gpu
@pytest.mark.parametrize("ty,typ", [("general", O.ISAI_GENERAL), ("lower", O.ISAI_LOWER), ("upper", O.ISAI_UPPER)])
def test_dropin_generate_isai_power_two(tmp_path, ty, typ):
    from test_dropin import write_problem
    print(f"case drop-in isai2_{ty}", flush=True)
    pat, A, b = synth.batch("stencil5", 16, n=100)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_problem(fin, pat, A, b)
    r = subprocess.run([DRIVER, "isai2_" + ty, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    mnnz = int(np.frombuffer(raw, np.int32, 1)[0])
    mrp = np.frombuffer(raw, np.int32, pat.n + 1, 4)
    mci = np.frombuffer(raw, np.int32, mnnz, 4 + 4 * (pat.n + 1))
    vals = np.frombuffer(raw, np.float64, 16 * mnnz, 4 + 4 * (pat.n + 1 + mnnz)).reshape(16, mnnz)
    wrp, wci = O.isai_pattern(pat.n, pat.row_ptrs, pat.col_idx, typ, 2)
    assert np.array_equal(mrp, wrp) and np.array_equal(mci, wci)
    for s in range(16):
        want, st = O.isai_on(pat, A[s], typ, wrp, wci)
        assert st == 0 and np.array_equal(vals[s], want)
