"""Oracle checks of the mixed-precision and warm-started ParILU(0) paths.

Each function is pinned by something other than its own formula: the
Jacobi ParILU(0) routine it extends, exact ILU(0) (the fixed point and the
fallback), the factor residual definition, numpy's dense solve and a CPU
recomputation of the true residual.
"""
import numpy as np
import pytest

import oracle as O
import synth


def relres(pat, A, b, x):
    D = pat.dense(A)
    return np.linalg.norm(b - D @ x) / np.linalg.norm(b)


def test_parilu_stop_fixed_sweeps_equals_jacobi():
    pat, A, _ = synth.batch("random", 4, n=144)
    for s in range(4):
        for sw in (0, 1, 3):
            F, st = O.parilu0_stop(pat, A[s], sw)
            Fj = O.parilu0(pat, A[s], sw)
            assert st == sw
            assert np.array_equal(F, Fj)


def test_parilu_stop_reaches_exact_ilu():
    pat, A, _ = synth.batch("random", 3, n=90)
    for s in range(3):
        Fe, _ = O.ilu0(pat, A[s])
        F, st = O.parilu0_stop(pat, A[s], 200, stop_tol=1e-14)
        assert 0 < st < 200
        assert np.allclose(F, Fe, rtol=1e-12, atol=0.0)
        # the kept iterate meets the tolerance by the residual's definition
        assert O.factor_residual(pat, A[s], F) <= 1e-14 * np.linalg.norm(A[s]) * 1.01


def test_parilu_stop_warm_start_at_fixed_point():
    pat, A, _ = synth.batch("random", 2, n=144)
    Fe, _ = O.ilu0(pat, A[0])
    F, st = O.parilu0_stop(pat, A[0], 10, stop_tol=1e-12, init=Fe)
    assert st == 0 and np.array_equal(F, Fe)


def test_parilu_stop_warm_start_only_entries_with_work():
    # entries without work keep a_ij whatever the guess says
    pat, A, _ = synth.batch("random", 1, n=64)
    junk = np.full(pat.nnz, 7.0)
    F, st = O.parilu0_stop(pat, A[0], 0, init=junk)
    order = O.entry_order(pat)
    rest = np.setdiff1d(np.arange(pat.nnz), order)
    assert np.array_equal(F[rest], A[0][rest])
    assert np.all(F[order] == 7.0)


def test_parilu_stop_fallback_is_exact_ilu():
    pat, A, _ = synth.batch("random", 2, n=144)
    Fe, _ = O.ilu0(pat, A[1])
    F, st = O.parilu0_stop(pat, A[1], 1, stop_tol=1e-14)
    assert st == -1 and np.array_equal(F, Fe)


def test_inplace_one_round_equals_jacobi():
    # with every entry with work in one round of 64, the in-place sweep reads
    # only the round's starting values: bit for bit a Jacobi sweep
    pat, A, _ = synth.batch("random", 3, n=16, density=0.2)
    assert 0 < len(O.sweep_order(pat)) <= 64
    for s in range(3):
        for sw in (1, 2, 5):
            Fi, si = O.parilu0_stop(pat, A[s], sw, inplace=True)
            Fj, sj = O.parilu0_stop(pat, A[s], sw)
            assert si == sj == sw and np.array_equal(Fi, Fj)


def test_inplace_two_rounds_differ_from_jacobi_and_reach_ilu():
    # more than one round: the later rounds see the earlier rounds' values, so
    # one in-place sweep differs from a Jacobi sweep; both share the fixed
    # point exact ILU(0)
    pat, A, _ = synth.batch("random", 2, n=144)
    assert len(O.sweep_order(pat)) > 64
    for s in range(2):
        Fi, _ = O.parilu0_stop(pat, A[s], 1, inplace=True)
        Fj, _ = O.parilu0_stop(pat, A[s], 1)
        assert not np.array_equal(Fi, Fj)
        Fe, _ = O.ilu0(pat, A[s])
        F, st = O.parilu0_stop(pat, A[s], 200, stop_tol=1e-14, inplace=True)
        assert 0 < st < 200
        assert np.allclose(F, Fe, rtol=1e-12, atol=0.0)


def test_inplace_round_order_contracts_faster():
    # in-place rounds in the plan's order (fewest products first) reach the
    # 1e-8 factor residual in fewer sweeps than Jacobi, from A and from the
    # base guess, and the kept iterate meets the tolerance by the residual's
    # definition (orc_factor_residual, an independent computation)
    pat, A, _ = synth.batch("random", 16, n=144)
    order = O.entry_order(pat)
    si, sj, sbi = [], [], []
    for s in range(16):
        F, k = O.parilu0_stop(pat, A[s], 40, stop_tol=1e-8, inplace=True)
        assert k > 0
        assert O.factor_residual(pat, A[s], F) <= 1e-8 * np.linalg.norm(A[s])
        si.append(k)
        sj.append(O.parilu0_stop(pat, A[s], 40, stop_tol=1e-8)[1])
        base = A[s].copy()
        rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
        low = rows > pat.col_idx
        base[low] = A[s][low] / A[s][pat.diag[pat.col_idx[low]]]
        sbi.append(O.parilu0_stop(pat, A[s], 40, stop_tol=1e-8, init=base, inplace=True)[1])
    assert np.mean(si) < np.mean(sj) - 1.5, (np.mean(si), np.mean(sj))
    assert np.mean(sbi) <= np.mean(si)


def test_inner_adapt_keeps_convergence_and_saves_inner_steps():
    # the adaptive inner tolerance (stop each FP32 inner solve once it takes
    # the FP64 residual below half the outer tolerance) keeps every system at
    # relres <= 1e-10 and spends fewer inner half steps than a fixed 1e-5
    pat, A, b = synth.batch("random", 96, n=144)
    res = {}
    for ad in (0.0, 0.5):
        o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, trsv_order=O.TRSV_AVAIL, inner_adapt=ad)
        x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precond=O.ILU0_ISAI, precision=O.PREC_FP32_IR,
                                                       o=o, threads=8)
        assert (cv == 1).all()
        assert max(relres(pat, A[s], b[s], x[s]) for s in range(A.shape[0])) <= 1e-10
        res[ad] = (it.mean(), rf.mean())
    assert res[0.5][0] < res[0.0][0]
    assert res[0.5][1] <= res[0.0][1] + 0.05


def test_entry_order_is_topological():
    # every entry's dependencies (ParILU pairs and the pivot) come at a lower
    # DAG level, so one pass in this order with level barriers is exact ILU(0)
    pat, _, _ = synth.batch("random", 1, n=90)
    order = O.entry_order(pat)
    rp, ci, diag = pat.row_ptrs, pat.col_idx, pat.diag
    pos = {(i, ci[p]): p for i in range(pat.n) for p in range(rp[i], rp[i + 1])}
    rank = {e: q for q, e in enumerate(order)}
    for e in order:
        i = int(np.searchsorted(rp, e, side="right") - 1)
        j = ci[e]
        for p in range(rp[i], diag[i]):
            k = ci[p]
            if k >= min(i, j) or (k, j) not in pos:
                continue
            for d in (p, pos[(k, j)]):
                assert d not in rank or rank[d] < rank[e]
        if j < i and diag[j] in rank:
            assert rank[diag[j]] < rank[e]


@pytest.mark.parametrize("kind,B,n", [("random", 24, 144), ("stencil5", 16, 100), ("random", 12, 300)])
def test_mixed_converges_to_fp64_accuracy(kind, B, n):
    pat, A, b = synth.batch(kind, B, n=n)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, trsv_order=O.TRSV_AVAIL)
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precision=O.PREC_FP32_IR, o=o,
                                                    threads=8)
    assert (cv == 1).all()
    for s in range(B):
        # independent CPU recomputation of the relative true residual
        rr = relres(pat, A[s], b[s], x[s])
        assert rr <= 1e-10
        # both recomputations carry rounding noise ~ eps * cond(A) ~ 1e-13
        assert tr[s] / np.linalg.norm(b[s]) == pytest.approx(rr, rel=1e-3, abs=1e-12)
        xd = np.linalg.solve(pat.dense(A[s]), b[s])
        assert np.linalg.norm(x[s] - xd) <= 1e-6 * np.linalg.norm(xd)
    # FP32 inner solves at 1e-5 cannot reach 1e-10 in one refinement step
    assert rf.min() >= 2 and rf.max() <= 6


def test_mixed_split_fp64_equals_merged_oracle():
    pat, A, b = synth.batch("random", 16, n=144)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1, trsv_order=O.TRSV_AVAIL)
    g = O.batch_solve_split(pat, A, b, precision=O.PREC_FP64, o=o, threads=4)
    m = O.batch_solve(pat, A, b, O.ILU0, o=o, threads=4)
    for u, v in zip(g[:5], m):
        assert np.array_equal(u, v)
    assert (g[5] == 0).all() and (g[6] == 0).all()


@pytest.mark.parametrize("pc", [O.ILU0_ISAI, O.PARILU0_ISAI])
def test_split_fp64_isai_equals_merged_oracle(pc):
    """The split driver's FP64 ISAI path (the persistent ISAI solve's
    oracle) is the batch driver's ISAI solver for the same factors, and its
    converged solutions agree with dense FP64 solves."""
    pat, A, b = synth.batch("random", 16, n=144)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE)
    g = O.batch_solve_split(pat, A, b, precond=pc, sweeps=5, precision=O.PREC_FP64, o=o, threads=4)
    m = O.batch_solve(pat, A, b, pc, sweeps=5, o=o, threads=4)
    for u, v in zip(g[:5], m):
        assert np.array_equal(u, v)
    assert (g[2] == 1).all()
    for s in range(4):
        xd = np.linalg.solve(pat.dense(A[s]), b[s])
        assert np.linalg.norm(g[0][s] - xd) <= 1e-6 * np.linalg.norm(xd)


def test_split_fp64_isai_zero_pivot():
    pat, A, b = synth.batch("random", 2, n=64)
    A = A.copy()
    A[1][pat.diag[0]] = 0.0
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE)
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precond=O.ILU0_ISAI, precision=O.PREC_FP64, o=o)
    assert cv[0] == 1 and cv[1] == O.ZERO_PIVOT
    assert (x[1] == 0).all() and it[1] == 0 and tr[1] == np.sqrt(O.dot(b[1], b[1]))


def test_mixed_special_cases():
    pat, A, b = synth.batch("random", 3, n=64)
    b = b.copy()
    b[0] = 0.0
    A = A.copy()
    A[2][pat.diag[5]] = 0.0       # zero pivot somewhere in the factorization
    A[2][pat.diag[0]] = 0.0
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE)
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precision=O.PREC_FP32_IR, o=o)
    assert cv[0] == 1 and it[0] == 0 and rf[0] == 0 and (x[0] == 0).all()
    assert cv[1] == 1
    # a zero leading pivot makes u_00 = 0: the FP32 factor check reports it
    assert cv[2] == O.ZERO_PIVOT and (x[2] == 0).all() and ir[2] == 0.0
    assert tr[2] == pytest.approx(np.linalg.norm(b[2]), rel=1e-15)


def test_mixed_exact_preconditioner_one_inner_step():
    # tridiagonal: ILU(0) is the exact LU, so each inner solve ends at its
    # first half step and refinement only fixes FP32 rounding
    n = 40
    rp, ci = [0], []
    for i in range(n):
        ci += [j for j in (i - 1, i, i + 1) if 0 <= j < n]
        rp.append(len(ci))
    pat = synth.Pattern(n, np.array(rp), np.array(ci))
    rng = np.random.default_rng(3)
    A = np.empty((4, pat.nnz))
    for s in range(4):
        v = -rng.uniform(0.5, 1.5, pat.nnz)
        for i in range(n):
            off = [p for p in range(pat.row_ptrs[i], pat.row_ptrs[i + 1]) if p != pat.diag[i]]
            v[pat.diag[i]] = np.abs(v[off]).sum() * 1.5
        A[s] = v
    b = rng.uniform(-1, 1, (4, n))
    o = O.opts(tol=1e-12, tol_mode=O.TOL_RELATIVE)
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precision=O.PREC_FP32_IR, o=o)
    assert (cv == 1).all()
    assert (it == 0.5 * rf).all()


def test_isai_f32_matches_fp64_isai_and_identity():
    # the FP32-storage sweep is the reference-order FP64 sweep (orc_isai) up
    # to FP32 rounding, and satisfies the defining identity (M L)(i, j) =
    # delta_ij on the pattern to FP32 accuracy
    pat, A, _ = synth.batch("random", 2, n=144)
    F, _ = O.ilu0(pat, A[0])
    M, st = O.isai_lu_f32(pat, F)
    assert st == 0
    ML, _ = O.isai(pat, F, O.ISAI_LOWER)
    MU, _ = O.isai(pat, F, O.ISAI_UPPER)
    rp, diag = pat.row_ptrs, pat.diag
    low = np.concatenate([np.arange(rp[i], diag[i]) for i in range(pat.n)])
    up = np.concatenate([np.arange(diag[i], rp[i + 1]) for i in range(pat.n)])
    assert np.allclose(M[low], ML[low], rtol=1e-4, atol=1e-6 * np.abs(ML[low]).max())
    assert np.allclose(M[up], MU[up], rtol=1e-4, atol=1e-6 * np.abs(MU[up]).max())
    Ld = np.eye(pat.n) + np.tril(pat.dense(F), -1)
    Ud = np.triu(pat.dense(F))
    MLd = np.eye(pat.n) + np.tril(pat.dense(M.astype(np.float64)), -1)
    MUd = np.triu(pat.dense(M.astype(np.float64)))
    mask = pat.dense(np.ones(pat.nnz)) != 0
    EL = (MLd @ Ld - np.eye(pat.n))[mask & np.tri(pat.n, dtype=bool)]
    EU = (MUd @ Ud - np.eye(pat.n))[mask & np.tri(pat.n, dtype=bool).T]
    assert np.abs(EL).max() < 1e-5 and np.abs(EU).max() < 1e-5


def test_isai_f32_exact_for_closed_patterns():
    # tridiagonal: L is bidiagonal and the power-1 ISAI of a bidiagonal
    # factor on its pattern reproduces the inverse's entries there
    n = 12
    rp, ci = [0], []
    for i in range(n):
        ci += [j for j in (i - 1, i, i + 1) if 0 <= j < n]
        rp.append(len(ci))
    pat = synth.Pattern(n, np.array(rp), np.array(ci))
    rng = np.random.default_rng(1)
    A = -rng.uniform(0.5, 1.5, pat.nnz)
    A[pat.diag] = 4.0
    F, _ = O.ilu0(pat, A)
    M, st = O.isai_lu_f32(pat, F)
    Linv = np.linalg.inv(np.eye(n) + np.tril(pat.dense(F), -1))
    Uinv = np.linalg.inv(np.triu(pat.dense(F)))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            want = Linv[i, j] if j < i else Uinv[i, j]
            assert M[p] == pytest.approx(want, rel=2e-7, abs=1e-7)


def test_isai_f32_zero_pivot():
    pat, A, _ = synth.batch("random", 1, n=40)
    F, _ = O.ilu0(pat, A[0])
    F[pat.diag[7]] = 0.0
    _, st = O.isai_lu_f32(pat, F)
    assert st == O.ZERO_PIVOT


def test_mixed_isai_converges():
    pat, A, b = synth.batch("random", 32, n=144)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE)
    for pc in (O.ILU0_ISAI, O.PARILU0_ISAI):
        x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precond=pc, precision=O.PREC_FP32_IR,
                                                        o=o, threads=8)
        assert (cv == 1).all()
        for s in range(0, 32, 5):
            assert relres(pat, A[s], b[s], x[s]) <= 1e-10


def test_split_general_isai():
    """The split driver's general ISAI (precond 4): FP64 = the batch driver's general-ISAI solver
    bit for bit; FP32 refinement over the FP32-rounded general ISAI reaches relres <= 1e-10
    (recomputed with dense products here) with 1 or 2 inner half steps per refinement on A = I"""
    pat, A, b = synth.batch("random", 16, n=144)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE)
    g = O.batch_solve_split(pat, A, b, precond=O.GENERAL_ISAI, precision=O.PREC_FP64, o=o, threads=4)
    m = O.batch_solve(pat, A, b, O.GENERAL_ISAI, o=o, threads=4)
    for u, v in zip(g[:5], m):
        assert np.array_equal(u, v)
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, A, b, precond=O.GENERAL_ISAI,
                                                    precision=O.PREC_FP32_IR, o=o, threads=4)
    assert (cv == 1).all() and (fs == 0).all() and (rf >= 1).all()
    for s in range(16):
        assert relres(pat, A[s], b[s], x[s]) <= 1e-10
    # A = I: the general ISAI is I, so every inner solve ends after one half step
    I = np.zeros_like(A[:2])
    I[:, pat.diag] = 1.0
    x, it, cv, ir, tr, rf, fs = O.batch_solve_split(pat, I, b[:2], precond=O.GENERAL_ISAI,
                                                    precision=O.PREC_FP32_IR, o=o)
    assert (cv == 1).all() and np.allclose(x, b[:2], rtol=1e-10, atol=0)
    assert (it == 0.5 * rf).all()
