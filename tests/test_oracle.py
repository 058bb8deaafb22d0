"""Pins the CPU oracle (oracle/oracle.c) against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, so each oracle
function is pinned by independent checks chosen so that a plausible slip (a
dropped term, a wrong index or sign, a transposed operand) fails one of
them: uniqueness of ILU(0) ((L U - A) on the pattern = 0), closed patterns
where ILU(0) is the dense LU (compared with scipy's LU), brute-force masked
Gaussian elimination, the ISAI defining identity, BiCGSTAB known answers and
dense solves (numpy.linalg.solve).
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
import synth


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not __import__("os").path.exists(O.LIB_PATH):
        import build as _b  # oracle/build.py
        _b.build()


def rand_pattern(n, density, seed):
    rng = np.random.default_rng(seed)
    M = rng.random((n, n)) < density
    M = M | M.T | np.eye(n, dtype=bool)
    rows, cols = np.nonzero(M)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return synth.Pattern(n, np.cumsum(rp), cols)


def dd_values(pat, seed, dom=1.5, asym=True):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, pat.nnz)
    off = pat.rows != pat.col_idx
    v[~off] = 0
    rs = np.zeros(pat.n)
    np.add.at(rs, pat.rows, np.abs(v))
    v[pat.diag] = rs * dom + 0.1
    return v


def dense_pattern(n):
    rows, cols = np.nonzero(np.ones((n, n), dtype=bool))
    return synth.Pattern(n, np.arange(0, n * n + 1, n), cols)


def lu_dense(pat, F):
    """unit-lower L and upper U as dense matrices from the in-place factor"""
    D = pat.dense(F)
    return np.tril(D, -1) + np.eye(pat.n), np.triu(D)


def masked_ge(A, P):
    """Textbook ILU(0): dense Gaussian elimination that drops fill outside P."""
    A = A.copy()
    n = A.shape[0]
    for k in range(n):
        for i in range(k + 1, n):
            if P[i, k]:
                A[i, k] /= A[k, k]
                A[i, k + 1:] -= np.where(P[i, k + 1:], A[i, k] * A[k, k + 1:], 0.0)
    return A


# ----------------------------------------------------------------- generators
def test_generator_counter_based_slices():
    pat, A, b = synth.batch("random", 16, n=144)
    _, A2, b2 = synth.batch("random", 6, first=5, n=144)
    assert np.array_equal(A[5:11], A2) and np.array_equal(b[5:11], b2)
    _, A3, _ = synth.batch("random", 16, n=144, chunk=3)
    assert np.array_equal(A, A3)


def test_generator_shapes_and_spec():
    pat, A, b = synth.batch("random", 32, n=144)
    assert abs(pat.nnz - 2000) < 150          # ~2000 nnz at density 0.09
    D = pat.dense(np.ones(pat.nnz)) != 0
    assert np.array_equal(D, D.T) and D.diagonal().all()
    off = pat.rows != pat.col_idx
    mags = np.abs(A[:, off])
    assert mags.min() >= 1e-6 and mags.max() <= 1e2
    rs = np.zeros((32, pat.n))
    np.add.at(rs, (slice(None), pat.rows), np.where(off, np.abs(A), 0))
    ratio = A[:, pat.diag] / rs
    assert ratio.min() >= 1.05 - 1e-12 and ratio.max() <= 2.0 + 1e-12
    assert np.abs(b).max() <= 1
    p5, A5, _ = synth.batch("stencil5", 4)
    assert p5.n == 100 and p5.nnz == 460
    assert A5[:, p5.rows != p5.col_idx].max() <= -0.5


# -------------------------------------------------------------------- kernels
def test_find_diag_bruteforce():
    pat = rand_pattern(30, 0.1, 1)
    d, missing = O.find_diag(pat.n, pat.row_ptrs, pat.col_idx)
    assert missing == 0
    for i in range(pat.n):
        assert pat.col_idx[d[i]] == i and pat.row_ptrs[i] <= d[i] < pat.row_ptrs[i + 1]
    # drop one diagonal
    keep = ~((pat.rows == 7) & (pat.col_idx == 7))
    rp = np.concatenate([[0], np.cumsum(np.bincount(pat.rows[keep], minlength=pat.n))])
    _, missing = O.find_diag(pat.n, rp, pat.col_idx[keep])
    assert missing == 1


@pytest.mark.parametrize("width", [1, 32, 64])
def test_dot_orders(width):
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(300), rng.standard_normal(300)
    assert O.dot(a, b, width) == pytest.approx(float(a @ b), rel=1e-13)
    # exactly representable partial sums -> exact
    ia = rng.integers(-8, 8, 300).astype(float)
    assert O.dot(ia, ia, width) == float(ia @ ia)


def test_dot_tree_order_matters_and_is_shfl_down():
    # width 2: (x0*y0 + x2*y2) + (x1*y1 + x3*y3) ; constructed so that the
    # association is visible in floating point
    a = np.array([1.0, 1e16, -1.0, -1e16])
    b = np.ones(4)
    assert O.dot(a, b, 1) == ((1.0 + 1e16) + -1.0) + -1e16
    assert O.dot(a, b, 2) == (1.0 + -1.0) + (1e16 + -1e16)


def test_spmv_vs_dense():
    pat = rand_pattern(40, 0.15, 2)
    v = dd_values(pat, 3)
    x = np.random.default_rng(4).standard_normal(pat.n)
    for w in (1, 32):
        assert np.allclose(O.spmv(pat, v, x, w), pat.dense(v) @ x, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n,density,seed", [(8, 0.3, 0), (8, 0.5, 1), (7, 0.4, 2), (20, 0.2, 3)])
def test_ilu0_bruteforce_masked_ge(n, density, seed):
    pat = rand_pattern(n, density, seed)
    v = dd_values(pat, seed + 10)
    F, st = O.ilu0(pat, v)
    assert st == 0
    P = pat.dense(np.ones(pat.nnz)) != 0
    G = masked_ge(pat.dense(v), P)
    assert np.allclose(pat.dense(F), np.where(P, G, 0), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("kind", ["random", "stencil5"])
def test_ilu0_uniqueness_residual(kind):
    pat, A, _ = synth.batch(kind, 4, n=144)
    for s in range(4):
        F, st = O.ilu0(pat, A[s])
        assert st == 0
        L, U = lu_dense(pat, F)
        P = pat.dense(np.ones(pat.nnz)) != 0
        R = np.where(P, L @ U - pat.dense(A[s]), 0)
        assert np.abs(R).max() <= 1e-13 * np.abs(A[s]).max()
        # the oracle's own residual agrees with the dense one
        assert O.factor_residual(pat, A[s], F) == pytest.approx(np.linalg.norm(R), abs=1e-13 * np.abs(A[s]).max())


def test_ilu0_closed_pattern_equals_dense_lu():
    n = 12
    pat = dense_pattern(n)
    rng = np.random.default_rng(5)
    Ad = rng.uniform(-1, 1, (n, n)) + np.diag(np.full(n, 3.0 * n))  # column dominant: no pivoting
    F, st = O.ilu0(pat, Ad.reshape(-1))
    Pm, Ls, Us = sla.lu(Ad)
    assert np.array_equal(Pm, np.eye(n))
    L, U = lu_dense(pat, F)
    assert np.allclose(L, Ls, rtol=1e-13, atol=1e-14) and np.allclose(U, Us, rtol=1e-13, atol=1e-13)
    # tridiagonal: ILU(0) is the exact LU as well
    T = np.diag(np.full(n, 4.0)) + np.diag(rng.uniform(-1, 1, n - 1), 1) + np.diag(rng.uniform(-1, 1, n - 1), -1)
    rows, cols = np.nonzero(T)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))])
    tp = synth.Pattern(n, rp, cols)
    F, _ = O.ilu0(tp, T[rows, cols])
    L, U = lu_dense(tp, F)
    assert np.allclose(L @ U, T, rtol=1e-14, atol=1e-14)


def test_ilu0_zero_pivot_detected():
    pat = rand_pattern(10, 0.3, 7)
    v = dd_values(pat, 8)
    v[pat.diag[0]] = 0.0
    _, st = O.ilu0(pat, v)
    assert st == O.ZERO_PIVOT


def test_parilu_fixed_point_is_ilu0():
    pat, A, _ = synth.batch("random", 2, n=144)
    F0, _ = O.ilu0(pat, A[0])
    # zero sweeps = the reference initial guess (L = tril(A)+I, U = triu(A))
    assert np.array_equal(O.parilu0(pat, A[0], 0), A[0])
    # Jacobi sweeps are exact after (dependency depth) sweeps
    F = O.parilu0(pat, A[0], pat.n)
    scale = np.abs(F0).max()
    assert np.abs(F - F0).max() <= 1e-12 * scale
    # and a fixed point of the sweep
    F2 = O.parilu0(pat, A[0], 1, init=F)
    assert np.abs(F2 - F).max() <= 1e-13 * scale
    assert O.factor_residual(pat, A[0], F) <= 1e-13 * np.abs(A[0]).max()
    # few sweeps: residual decreases from the initial guess
    r0 = O.factor_residual(pat, A[0], A[0])
    r5 = O.factor_residual(pat, A[0], O.parilu0(pat, A[0], 5))
    assert r5 < r0


def test_parilu_one_sweep_dense_formula():
    # one Jacobi sweep from (L0,U0): L1 = (A - strictly-lower part of L0 U0
    # restricted) ... checked entrywise against dense products of the guess
    pat = rand_pattern(10, 0.4, 11)
    A = dd_values(pat, 12)
    F1 = O.parilu0(pat, A, 1)
    L0, U0 = lu_dense(pat, A)
    P = pat.dense(np.ones(pat.nnz)) != 0
    Ad = pat.dense(A)
    for i, j in zip(pat.rows, pat.col_idx):
        m = min(i, j)
        s = L0[i, :m] @ U0[:m, j]
        want = (Ad[i, j] - s) / U0[j, j] if i > j else Ad[i, j] - s
        got = pat.dense(F1)[i, j]
        assert got == pytest.approx(want, rel=1e-13, abs=1e-14)
    assert P.any()


@pytest.mark.parametrize("typ", [O.ISAI_LOWER, O.ISAI_UPPER, O.ISAI_GENERAL])
def test_isai_defining_identity(typ):
    pat, A, _ = synth.batch("random", 1, n=144)
    F, _ = O.ilu0(pat, A[0])
    src = A[0] if typ == O.ISAI_GENERAL else F
    M, st = O.isai(pat, src, typ)
    assert st == 0
    if typ == O.ISAI_LOWER:
        Md = np.tril(pat.dense(M))
        Xd = np.tril(pat.dense(F), -1) + np.eye(pat.n)
        Pm = np.tril(pat.dense(np.ones(pat.nnz)) != 0)
    elif typ == O.ISAI_UPPER:
        Md = np.triu(pat.dense(M))
        Xd = np.triu(pat.dense(F))
        Pm = np.triu(pat.dense(np.ones(pat.nnz)) != 0)
    else:
        Md = pat.dense(M)
        Xd = pat.dense(A[0])
        Pm = pat.dense(np.ones(pat.nnz)) != 0
    R = (Md @ Xd - np.eye(pat.n))[Pm]
    assert np.abs(R).max() <= 1e-9 * max(1.0, np.abs(Md).max())


def test_isai_exact_inverse_cases():
    n = 9
    # diagonal pattern: ISAI = exact inverse
    pat = synth.Pattern(n, np.arange(n + 1), np.arange(n))
    d = np.linspace(1, 3, n)
    M, _ = O.isai(pat, d, O.ISAI_GENERAL)
    assert np.allclose(M, 1 / d, rtol=1e-15)
    # full lower triangle pattern: ISAI(L) = exact L^{-1}
    dp = dense_pattern(n)
    rng = np.random.default_rng(0)
    Fv = rng.uniform(-1, 1, n * n)
    M, _ = O.isai(dp, Fv, O.ISAI_LOWER)
    L = np.tril(dp.dense(Fv), -1) + np.eye(n)
    assert np.allclose(np.tril(dp.dense(M)), np.linalg.inv(L), rtol=1e-12, atol=1e-12)
    U = np.triu(dp.dense(Fv)) + np.eye(n) * 4
    M, _ = O.isai(dp, U[np.nonzero(np.ones((n, n)))], O.ISAI_UPPER)
    assert np.allclose(np.triu(dp.dense(M)), np.linalg.inv(U), rtol=1e-12, atol=1e-12)
    # general with pivoting on a dense pattern: exact inverse
    G = rng.uniform(-1, 1, (n, n))
    M, _ = O.isai(dp, G.reshape(-1), O.ISAI_GENERAL)
    assert np.allclose(dp.dense(M), np.linalg.inv(G), rtol=1e-10, atol=1e-10)


# ------------------------------------------------------------------ BiCGSTAB
def test_bicgstab_known_answers():
    n = 20
    ip = synth.Pattern(n, np.arange(n + 1), np.arange(n))
    b = np.linspace(-1, 1, n) + 0.01
    # A = I, no preconditioner: converges at the first half step
    x, it, cv, ir, tr = O.bicgstab(ip, np.ones(n), b, O.PC_NONE)
    assert it == 0.5 and cv == 1 and np.array_equal(x, b)
    # b = 0: x = 0, iteration 0, flag 1
    x, it, cv, ir, tr = O.bicgstab(ip, np.ones(n), np.zeros(n), O.PC_NONE)
    assert it == 0 and cv == 1 and not x.any() and tr == 0
    # ||b|| below the absolute tolerance
    x, it, cv, ir, tr = O.bicgstab(ip, np.ones(n), np.full(n, 1e-14), O.PC_NONE)
    assert it == 0 and cv == 1
    # exact preconditioner (closed pattern): converges at iteration 0.5
    dp = dense_pattern(8)
    Ad = np.random.default_rng(1).uniform(-1, 1, (8, 8)) + np.eye(8) * 10
    F, _ = O.ilu0(dp, Ad.reshape(-1))
    x, it, cv, ir, tr = O.bicgstab(dp, Ad.reshape(-1), np.ones(8), O.PC_LU, F)
    assert it == 0.5 and cv == 1
    assert np.allclose(x, np.linalg.solve(Ad, np.ones(8)), rtol=1e-13)


@pytest.mark.parametrize("kind,pc", [("stencil5", O.ILU0), ("random", O.ILU0), ("random", O.PARILU0),
                                     ("random", O.ILU0_ISAI), ("random", O.GENERAL_ISAI)])
def test_bicgstab_matches_dense_solve(kind, pc):
    pat, A, b = synth.batch(kind, 8, n=144)
    o = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1)
    X, IT, CV, IR, TR = O.batch_solve(pat, A, b, pc, sweeps=5, o=o)
    assert (CV == 1).all()
    for s in range(8):
        Ad = pat.dense(A[s])
        assert np.linalg.norm(b[s] - Ad @ X[s]) / np.linalg.norm(b[s]) < 1e-10
        # the residual at 1e-10 relative carries rounding noise of its own
        assert TR[s] == pytest.approx(np.linalg.norm(b[s] - Ad @ X[s]), rel=1e-3, abs=1e-16)
        xd = np.linalg.solve(Ad, b[s])
        assert np.linalg.norm(X[s] - xd) <= 1e-4 * np.linalg.norm(xd)  # cond(A) ~ 1e6


def test_reference_vs_wave64_order_iterations_within_one():
    """Reduction order (warp-32 reference vs wave64 kernels) moves the
    iteration count by at most one half step pair."""
    for kind in ("stencil5", "random"):
        pat, A, b = synth.batch(kind, 64, n=144)
        _, it32, cv32, _, _ = O.batch_solve(pat, A, b, O.ILU0, o=O.reference_opts())
        _, it64, cv64, _, _ = O.batch_solve(pat, A, b, O.ILU0, o=O.opts())
        assert (cv32 == 1).all() and (cv64 == 1).all()
        assert np.abs(it32 - it64).max() <= 1.0


def test_bicgstab_failure_modes():
    pat, A, b = synth.batch("random", 4, n=144)
    # max_iter = 1 -> flag -1 after at most one full iteration
    _, it, cv, _, _ = O.batch_solve(pat, A, b, O.NONE, o=O.opts(max_iter=1))
    assert (cv == -1).all() and (it <= 1).all()
    # NaN in A: non-finite detected, no hang
    A2 = A.copy()
    A2[1, 5] = np.nan
    _, it, cv, _, _ = O.batch_solve(pat, A2, b, O.ILU0)
    assert cv[1] == O.NONFINITE and cv[0] == 1
    # zero pivot
    A3 = A.copy()
    A3[2, pat.diag[0]] = 0.0
    x, it, cv, _, tr = O.batch_solve(pat, A3, b, O.ILU0)
    assert cv[2] == O.ZERO_PIVOT and not x[2].any() and tr[2] == pytest.approx(np.linalg.norm(b[2]))


def test_threads_do_not_change_bits():
    pat, A, b = synth.batch("random", 24, n=144)
    r1 = O.batch_solve(pat, A, b, O.ILU0, threads=1)
    r4 = O.batch_solve(pat, A, b, O.ILU0, threads=4)
    for u, v in zip(r1, r4):
        assert np.array_equal(u, v)


def test_availability_order_solve_is_the_same_solve():
    """The persistent kernel's row-sum order (entries by level of their
    column) is another evaluation order of the same exact triangular solves."""
    pat, A, b = synth.batch("random", 4, n=144)
    for s in range(4):
        F, _ = O.ilu0(pat, A[s])
        x1 = O.lu_solve(pat, F, b[s])
        x2 = O.lu_solve_avail(pat, F, b[s])
        L, U = lu_dense(pat, F)
        xd = np.linalg.solve(U, np.linalg.solve(L, b[s]))
        assert np.allclose(x1, xd, rtol=1e-9, atol=0) and np.allclose(x2, xd, rtol=1e-9, atol=0)
        assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x1).max()
    ll, lu = O.trsv_levels(pat)
    for i in range(pat.n):
        cols = pat.col_idx[pat.row_ptrs[i]:pat.diag[i]]
        assert ll[i] == (1 + ll[cols].max() if cols.size else 0)


def test_availability_order_iterations_within_one():
    for kind in ("stencil5", "random"):
        pat, A, b = synth.batch(kind, 64, n=144)
        _, it0, cv0, _, _ = O.batch_solve(pat, A, b, O.ILU0, o=O.opts())
        _, it1, cv1, _, _ = O.batch_solve(pat, A, b, O.ILU0, o=O.opts(trsv_order=O.TRSV_AVAIL))
        assert (cv0 == 1).all() and (cv1 == 1).all()
        assert np.abs(it0 - it1).max() <= 1.0
