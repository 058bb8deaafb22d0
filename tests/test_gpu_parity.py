"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

The kernels and the oracle use the same operation order (explicit fma,
wave64 dot order, CSR-order row sums), so results are compared bit for bit:
factors, solutions, iteration counts, flags and residual norms.  Sizes are
chosen so the oracle finishes in seconds; the full-size batch is covered by
size-independent properties (every system converged, true relres <= tol,
spot checks against the oracle).
"""
import numpy as np
import pytest

import bilu
import oracle as O
import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev(), dtype=dtype)


def run_gpu(pat, A, b, **kw):
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    x = torch.empty_like(db)
    it, cv, ir, tr = bilu.solve(P, dA, db, x, bilu.opts(**kw))
    torch.cuda.synchronize()
    return x.cpu().numpy(), it.cpu().numpy(), cv.cpu().numpy(), ir.cpu().numpy(), tr.cpu().numpy()


ORACLE_PC = {bilu.ILU0: O.ILU0, bilu.PARILU0: O.PARILU0, bilu.ILU0_ISAI: O.ILU0_ISAI,
             bilu.PARILU0_ISAI: O.PARILU0_ISAI, bilu.GENERAL_ISAI: O.GENERAL_ISAI, bilu.NONE: O.NONE}


def uses_persistent(pat, precond=bilu.ILU0, kernel=bilu.KERNEL_AUTO, **_):
    """Whether bilu_solve runs the persistent kernel (LU preconditioners, pattern
    within its register budget), whose triangular-solve row sums run in
    availability order."""
    if precond not in (bilu.ILU0, bilu.PARILU0) or kernel == bilu.KERNEL_WAVE:
        return False
    return pat.n <= 192 and pat.nnz <= 48 * 64


def run_oracle(pat, A, b, precond=bilu.ILU0, parilu_sweeps=5, tol=1e-11, tol_mode=0, max_iter=500,
               true_res_restart=0, max_restarts=3, kernel=bilu.KERNEL_AUTO, **_):
    order = O.TRSV_AVAIL if uses_persistent(pat, precond, kernel) else O.TRSV_CSR
    o = O.opts(tol=tol, tol_mode=tol_mode, max_iter=max_iter, true_res_restart=true_res_restart,
               max_restarts=max_restarts, trsv_order=order)
    return O.batch_solve(pat, A, b, ORACLE_PC[precond], sweeps=parilu_sweeps, o=o, threads=8)


def assert_same(g, o, what=""):
    names = ["x", "iters", "conv", "iter_res", "true_res"]
    for name, u, v in zip(names, g, o):
        assert np.array_equal(u, v), f"{what}: {name} differs (max |diff| {np.nanmax(np.abs(u - v))})"


CASES = [("stencil5", 64, 100), ("random", 256, 144), ("random", 64, 32), ("random", 64, 90),
         ("random", 32, 300)]


@pytest.mark.parametrize("kernel", [bilu.KERNEL_WAVE, bilu.KERNEL_PERSISTENT])
@pytest.mark.parametrize("kind,B,n", CASES)
def test_merged_ilu0_bitwise(kind, B, n, kernel):
    print(f"case merged ilu0 {kind} B={B} n={n} kernel={kernel}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    if kernel == bilu.KERNEL_PERSISTENT and n == 300:
        with pytest.raises(bilu.BiluError):   # 65 SpMV slots per lane: beyond the register budget
            run_gpu(pat, A, b, kernel=kernel)
        return
    g = run_gpu(pat, A, b, kernel=kernel)
    o = run_oracle(pat, A, b, kernel=kernel)
    assert (o[2] == 1).all()
    assert_same(g, o, f"{kind} n={n}")


@pytest.mark.parametrize("kernel", [bilu.KERNEL_WAVE, bilu.KERNEL_PERSISTENT])
def test_persistent_parilu_and_segregated_bitwise(kernel):
    pat, A, b = synth.batch("random", 200, n=144)
    for kw in (dict(precond=bilu.PARILU0, parilu_sweeps=3),
               dict(precond=bilu.ILU0, merge=bilu.SEGREGATED),
               dict(precond=bilu.PARILU0, parilu_sweeps=5, merge=bilu.SEGREGATED),
               dict(precond=bilu.ILU0, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, true_res_restart=1)):
        print(f"case {kw} kernel={kernel}", flush=True)
        g = run_gpu(pat, A, b, kernel=kernel, **kw)
        o = run_oracle(pat, A, b, kernel=kernel, **kw)
        assert_same(g, o, f"{kw} kernel={kernel}")


@pytest.mark.parametrize("precond", [bilu.PARILU0, bilu.ILU0_ISAI, bilu.PARILU0_ISAI,
                                     bilu.GENERAL_ISAI, bilu.NONE])
def test_precond_variants_bitwise(precond):
    print(f"case precond {precond}", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    kw = dict(precond=precond, parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"precond {precond}")


@pytest.mark.parametrize("precond", [bilu.ILU0_ISAI, bilu.PARILU0_ISAI, bilu.GENERAL_ISAI])
@pytest.mark.parametrize("kind,B,n", [("stencil5", 64, 100), ("random", 128, 144), ("random", 64, 32),
                                      ("random", 64, 90)])
def test_persistent_isai_bitwise(precond, kind, B, n):
    """FP64 ISAI solvers on the persistent kernel (generation launch, then the
    solve applying ISAI(L), ISAI(U) or the general ISAI as SpMVs): bitwise the
    oracle and the one-wave-per-system kernel."""
    print(f"case persistent isai precond={precond} {kind} B={B} n={n}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    kw = dict(precond=precond, parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
    g = run_gpu(pat, A, b, kernel=bilu.KERNEL_PERSISTENT, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert (o[2] == 1).all()
    assert_same(g, o, f"persistent isai {precond} {kind} n={n}")
    w = run_gpu(pat, A, b, kernel=bilu.KERNEL_WAVE, **kw)
    assert_same(g, w, f"persistent vs wave isai {precond} {kind} n={n}")


def test_relative_tol_with_restart_meets_1e10():
    pat, A, b = synth.batch("random", 512, n=144)
    kw = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, true_res_restart=1)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, "relative+restart")
    relres = g[4] / np.linalg.norm(b, axis=1)
    assert (g[2] == 1).all() and relres.max() < 1e-10


def test_segregated_equals_merged():
    pat, A, b = synth.batch("random", 128, n=144)
    for pc in (bilu.ILU0, bilu.PARILU0):
        m = run_gpu(pat, A, b, precond=pc)
        s = run_gpu(pat, A, b, precond=pc, merge=bilu.SEGREGATED)
        assert_same(m, s, f"segregated pc={pc}")


def test_factor_entry_points_bitwise():
    pat, A, _ = synth.batch("random", 64, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA = to_dev(A)
    F = torch.empty_like(dA)
    st = torch.empty(64, dtype=torch.int32, device=dev())
    bilu.ilu0(P, dA, F, st)
    torch.cuda.synchronize()
    Fo = np.stack([O.ilu0(pat, A[s])[0] for s in range(64)])
    assert np.array_equal(F.cpu().numpy(), Fo) and not st.cpu().numpy().any()
    # factor residual (reduction order differs: tolerance)
    res = torch.empty(64, dtype=torch.float64, device=dev())
    bilu.factor_residual(P, dA, F, res)
    ro = np.array([O.factor_residual(pat, A[s], Fo[s]) for s in range(64)])
    assert np.allclose(res.cpu().numpy(), ro, rtol=1e-6, atol=1e-13 * np.abs(A).max())
    # ParILU sweeps: Jacobi, bitwise
    for sweeps in (0, 1, 5, 40):
        Fp = torch.empty_like(dA)
        bilu.parilu0(P, dA, Fp, sweeps)
        torch.cuda.synchronize()
        Fpo = np.stack([O.parilu0(pat, A[s], sweeps) for s in range(64)])
        assert np.array_equal(Fp.cpu().numpy(), Fpo), f"parilu sweeps={sweeps}"
    # reference L/U split
    lrp, lci, urp, uci = P.lu_split()
    inf = P.info()
    L = torch.empty(64, inf.l_nnz, dtype=torch.float64, device=dev())
    U = torch.empty(64, inf.u_nnz, dtype=torch.float64, device=dev())
    bilu.split_lu(P, F, L, U)
    torch.cuda.synchronize()
    Ln, Un = L.cpu().numpy(), U.cpu().numpy()
    for s in (0, 63):
        Fd = pat.dense(Fo[s])
        Ld = np.zeros((pat.n, pat.n))
        Ld[np.repeat(np.arange(pat.n), np.diff(lrp)), lci] = Ln[s]
        Ud = np.zeros((pat.n, pat.n))
        Ud[np.repeat(np.arange(pat.n), np.diff(urp)), uci] = Un[s]
        assert np.array_equal(Ld, np.tril(Fd, -1) + np.eye(pat.n))
        assert np.array_equal(Ud, np.triu(Fd))


@pytest.mark.parametrize("typ", [bilu.ISAI_LOWER, bilu.ISAI_UPPER, bilu.ISAI_GENERAL])
def test_isai_bitwise(typ):
    pat, A, _ = synth.batch("random", 32, n=144)
    Fo = np.stack([O.ilu0(pat, A[s])[0] for s in range(32)])
    src = A if typ == bilu.ISAI_GENERAL else Fo
    P = bilu.Pattern.from_synth(pat)
    out = torch.zeros(32, pat.nnz, dtype=torch.float64, device=dev())
    st = torch.empty(32, dtype=torch.int32, device=dev())
    bilu.isai(P, to_dev(src), typ, out, st)
    torch.cuda.synchronize()
    want = np.stack([O.isai(pat, src[s], typ)[0] for s in range(32)])
    assert np.array_equal(out.cpu().numpy(), want) and not st.cpu().numpy().any()


@pytest.mark.parametrize("n", [32, 144, 300])
def test_isai_general_pivoting_bitwise(n):
    """general ISAI on blocks without diagonal dominance, so Gauss-Jordan exchanges rows: values
    +-1..9 (equal magnitudes tie and the pivot tree's order decides; at n = 144 the tree picks
    another row than the first largest in about 1% of the steps) and standard normal"""
    print(f"case isai general pivoting n={n}", flush=True)
    pat, _, _ = synth.batch("random", 1, n=n)
    rng = np.random.default_rng(7)
    B = 32
    ints = np.concatenate([np.arange(1.0, 10.0), -np.arange(1.0, 10.0)])
    A = np.concatenate([rng.choice(ints, size=(B // 2, pat.nnz)),
                        rng.standard_normal((B // 2, pat.nnz))])
    P = bilu.Pattern.from_synth(pat)
    out = torch.zeros(B, pat.nnz, dtype=torch.float64, device=dev())
    st = torch.empty(B, dtype=torch.int32, device=dev())
    bilu.isai(P, to_dev(A), bilu.ISAI_GENERAL, out, st)
    torch.cuda.synchronize()
    res = [O.isai(pat, A[s], O.ISAI_GENERAL) for s in range(B)]
    want = np.stack([r[0] for r in res])
    ost = np.array([r[1] for r in res])
    gst = st.cpu().numpy()
    assert np.array_equal(gst != 0, ost != 0)
    ok = ost == 0
    assert ok.sum() >= B - 4
    assert np.array_equal(out.cpu().numpy()[ok].view(np.int64), want[ok].view(np.int64))


def test_true_residual_kernel():
    pat, A, b = synth.batch("random", 16, n=144)
    x = np.random.default_rng(0).standard_normal(b.shape)
    P = bilu.Pattern.from_synth(pat)
    res = torch.empty(16, dtype=torch.float64, device=dev())
    bilu.true_residual(P, to_dev(A), to_dev(b), to_dev(x), res)
    torch.cuda.synchronize()
    want = np.array([np.sqrt(O.dot(b[s] - O.spmv(pat, A[s], x[s]), b[s] - O.spmv(pat, A[s], x[s])))
                     for s in range(16)])
    assert np.array_equal(res.cpu().numpy(), want)


@pytest.mark.parametrize("precond", [bilu.ILU0_ISAI, bilu.PARILU0_ISAI, bilu.GENERAL_ISAI])
def test_wave_kernel_loaded_isai_zero_pivot(precond):
    """one-wave solve kernel with ISAI values from their own generation launches: a zero pivot
    (a zero diagonal of A: the first pivot for ILU(0), a singular block for the general ISAI) flags the
    system (-2, x = 0) as the oracle does; the other systems bitwise the oracle"""
    print(f"case wave loaded isai zero pivot precond={precond}", flush=True)
    pat, A, b = synth.batch("random", 8, n=144)
    A = A.copy()
    if precond == bilu.GENERAL_ISAI:
        A[2, pat.row_ptrs[0]:pat.row_ptrs[1]] = 0.0      # row 0 of A zero
    else:
        A[2, pat.diag[0]] = 0.0                          # u_00 = 0
    kw = dict(precond=precond, parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
    g = run_gpu(pat, A, b, kernel=bilu.KERNEL_WAVE, **kw)
    o = run_oracle(pat, A, b, kernel=bilu.KERNEL_WAVE, **kw)
    assert g[2][2] == bilu.ZERO_PIVOT and not g[0][2].any()
    assert_same(g, o, f"wave loaded isai zero pivot {precond}")


def test_edge_cases_no_hang():
    pat, A, b = synth.batch("random", 8, n=144)
    A = A.copy()
    b = b.copy()
    b[0] = 0.0                      # b = 0 -> x = 0, iter 0, flag 1
    A[1, 7] = np.nan                # non-finite -> flag -3, no hang
    A[2, pat.diag[0]] = 0.0         # zero pivot -> flag -2
    b[3] = 1e-14                    # ||b|| < ATOL -> iter 0, flag 1
    g = run_gpu(pat, A, b)
    o = run_oracle(pat, A, b)
    x, it, cv, ir, tr = g
    assert cv[0] == 1 and it[0] == 0 and not x[0].any()
    assert cv[1] == bilu.NONFINITE
    assert cv[2] == bilu.ZERO_PIVOT and not x[2].any()
    assert cv[3] == 1 and it[3] == 0
    assert (cv[4:] == 1).all()
    assert np.array_equal(cv, o[2]) and np.array_equal(it, o[1])
    ok = np.array([0, 2, 3, 4, 5, 6, 7])
    assert np.array_equal(x[ok], o[0][ok])
    # max_iter = 1
    g = run_gpu(pat, A[4:], b[4:], max_iter=1, precond=bilu.NONE)
    assert (g[2] == -1).all() and (g[1] <= 1).all()


def test_empty_batch_and_bad_args():
    pat, A, b = synth.batch("random", 1, n=144)
    P = bilu.Pattern.from_synth(pat)
    e = torch.empty(0, pat.nnz, dtype=torch.float64, device=dev())
    eb = torch.empty(0, pat.n, dtype=torch.float64, device=dev())
    bilu.solve(P, e, eb, eb.clone())
    with pytest.raises(bilu.BiluError):
        bilu.Pattern(3, [0, 1, 2, 3], [0, 2, 1])        # missing diagonal in rows 1, 2
    with pytest.raises(bilu.BiluError):
        bilu.Pattern(2, [0, 2, 3], [1, 0, 1])           # unsorted row
    big = synth.random_pattern(513, 0.01)                 # beyond the 512-row cap
    with pytest.raises(bilu.BiluError):
        Pb = bilu.Pattern.from_synth(big)
        bilu.solve(Pb, torch.zeros(1, big.nnz, dtype=torch.float64, device=dev()),
                   torch.ones(1, big.n, dtype=torch.float64, device=dev()),
                   torch.zeros(1, big.n, dtype=torch.float64, device=dev()))


def test_full_size_batch_properties():
    """65536 systems of the n=144 pattern: all converge to relres < 1e-10;
    a sample is bitwise equal to the oracle."""
    B = 65536
    pat, A, b = synth.batch("random", B, n=144)
    kw = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, true_res_restart=1)
    x, it, cv, ir, tr = run_gpu(pat, A, b, **kw)
    relres = tr / np.linalg.norm(b, axis=1)
    assert (cv == 1).all() and relres.max() < 1e-10
    idx = np.random.default_rng(1).choice(B, 256, replace=False)
    o = run_oracle(pat, A[idx], b[idx], **kw)
    assert_same((x[idx], it[idx], cv[idx], ir[idx], tr[idx]), o, "full-size sample")
