"""GPU parity of the split factor + solve paths against the oracle:
FP32-inner / FP64-refinement solves and warm-started, residual-stopped
ParILU(0).  Same operation order on both sides, so results are compared bit
for bit (solutions, iteration and refinement counts, flags, residuals,
factor status); the full-size batch is covered by size-independent
properties.
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


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def run_gpu(pat, A, b, F_init=None, **kw):
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    x = torch.empty_like(db)
    out = bilu.solve_ex(P, dA, db, x, bilu.opts(**kw), F_init=None if F_init is None else to_dev(F_init))
    torch.cuda.synchronize()
    return (x.cpu().numpy(),) + tuple(t.cpu().numpy() for t in out)


def run_oracle(pat, A, b, precond=bilu.ILU0, parilu_sweeps=5, tol=1e-10, tol_mode=1,
               precision=0, inner_tol=1e-5, max_refine=10, parilu_stop_tol=0.0, init=None,
               true_res_restart=0, max_iter=500, parilu_order=bilu.PARILU_JACOBI, **_):
    o = O.opts(tol=tol, tol_mode=tol_mode, max_iter=max_iter, true_res_restart=true_res_restart,
               trsv_order=O.TRSV_AVAIL)
    pc = {bilu.ILU0: O.ILU0, bilu.PARILU0: O.PARILU0, bilu.ILU0_ISAI: O.ILU0_ISAI,
          bilu.PARILU0_ISAI: O.PARILU0_ISAI, bilu.GENERAL_ISAI: O.GENERAL_ISAI}[precond]
    return O.batch_solve_split(pat, A, b, precond=pc,
                               sweeps=parilu_sweeps, init=init, stop_tol=parilu_stop_tol,
                               precision=precision, inner_tol=inner_tol, max_refine=max_refine,
                               o=o, threads=8, inplace=parilu_order == bilu.PARILU_INPLACE)


NAMES = ["x", "iters", "conv", "iter_res", "true_res", "refine", "factor_status"]


def assert_same(g, o, what):
    for name, u, v in zip(NAMES, g, o):
        assert np.array_equal(u, v), f"{what}: {name} differs (max |diff| {np.nanmax(np.abs(u - v))})"


@pytest.mark.parametrize("kind,B,n", [("random", 256, 144), ("stencil5", 64, 100), ("random", 64, 32),
                                      ("random", 64, 90)])
@pytest.mark.parametrize("inner_tol", [1e-5, 1e-3])
def test_mixed_bitwise(kind, B, n, inner_tol):
    print(f"case mixed {kind} B={B} n={n} inner_tol={inner_tol}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    kw = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR, inner_tol=inner_tol)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"mixed {kind} n={n}")
    assert (o[2] == 1).all()


@pytest.mark.parametrize("kind,B,n", [("random", 256, 144), ("stencil5", 64, 100), ("random", 64, 32),
                                      ("random", 64, 90)])
@pytest.mark.parametrize("precond", [bilu.ILU0_ISAI, bilu.PARILU0_ISAI])
def test_mixed_isai_bitwise(kind, B, n, precond):
    print(f"case mixed isai {kind} B={B} n={n} precond={precond}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    kw = dict(precond=precond, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"mixed isai {kind} n={n}")
    assert (o[2] == 1).all()


def test_mixed_isai_warm_parilu_bitwise():
    print("case mixed isai warm parilu", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    F = np.stack([O.ilu0(pat, A[s])[0] for s in range(A.shape[0])])
    init = F * (1.0 + 1e-3 * np.random.default_rng(7).standard_normal(F.shape))
    kw = dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    assert_same(run_gpu(pat, A, b, F_init=init, **kw), run_oracle(pat, A, b, init=init, **kw),
                "mixed isai warm parilu")


def test_mixed_isai_zero_pivot():
    print("case mixed isai zero pivot", flush=True)
    pat, A, b = synth.batch("random", 4, n=64)
    A = A.copy()
    A[1][pat.diag[0]] = 0.0
    kw = dict(precond=bilu.ILU0_ISAI, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    assert_same(g, run_oracle(pat, A, b, **kw), "mixed isai zero pivot")
    assert g[2][1] == bilu.ZERO_PIVOT and (g[2][[0, 2, 3]] == 1).all()


def test_mixed_parilu_bitwise():
    print("case mixed parilu 5 sweeps", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    kw = dict(precond=bilu.PARILU0, parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE,
              precision=bilu.PRECISION_FP32_IR)
    assert_same(run_gpu(pat, A, b, **kw), run_oracle(pat, A, b, **kw), "mixed parilu")


@pytest.mark.parametrize("precision", [bilu.PRECISION_FP64, bilu.PRECISION_FP32_IR])
def test_warm_parilu_stop_bitwise(precision):
    print(f"case warm parilu precision={precision}", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    # warm start: exact factors with a relative perturbation (a stand-in for a learned guess)
    F = np.stack([O.ilu0(pat, A[s])[0] for s in range(A.shape[0])])
    rng = np.random.default_rng(5)
    init = F * (1.0 + 1e-3 * rng.standard_normal(F.shape))
    kw = dict(precond=bilu.PARILU0, parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=precision, true_res_restart=1)
    g = run_gpu(pat, A, b, F_init=init, **kw)
    o = run_oracle(pat, A, b, init=init, **kw)
    assert_same(g, o, "warm parilu")
    assert (o[6] > 0).all()           # sweeps were needed and the tolerance was met


def test_parilu_stop_fallback_bitwise():
    print("case parilu fallback", flush=True)
    pat, A, b = synth.batch("random", 64, n=144)
    kw = dict(precond=bilu.PARILU0, parilu_sweeps=2, parilu_stop_tol=1e-12, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, "parilu fallback")
    assert (g[6] == -1).all()


def test_mixed_special_cases():
    print("case mixed special", flush=True)
    pat, A, b = synth.batch("random", 3, n=64)
    b = b.copy(); A = A.copy()
    b[0] = 0.0
    A[2][pat.diag[0]] = 0.0
    kw = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    assert_same(g, run_oracle(pat, A, b, **kw), "mixed special")
    assert g[2][2] == bilu.ZERO_PIVOT


@pytest.mark.parametrize("kind,B,n", [("random", 256, 144), ("stencil5", 64, 100), ("random", 64, 32),
                                      ("random", 64, 300)])
def test_mixed_general_isai_bitwise(kind, B, n):
    """FP32 general ISAI of A (generated in FP64 with the reference pivot tree, stored FP32, applied as
    one SpMV over A's slot records) inside FP64 refinement: bitwise the oracle, every system to 1e-10"""
    print(f"case mixed general isai {kind} B={B} n={n}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    kw = dict(precond=bilu.GENERAL_ISAI, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"mixed general isai {kind} n={n}")
    assert (o[2] == 1).all()
    relres = np.array([np.linalg.norm(b[s] - pat.dense(A[s]) @ g[0][s]) / np.linalg.norm(b[s])
                       for s in range(0, B, max(1, B // 16))])
    assert relres.max() <= 1e-10 * 1.0001


def test_mixed_general_isai_zero_pivot():
    print("case mixed general isai zero pivot", flush=True)
    pat, A, b = synth.batch("random", 8, n=32)
    A = A.copy()
    row0 = slice(pat.row_ptrs[0], pat.row_ptrs[1])
    A[3, row0] = 0.0   # row 0 of A zero: every block holding it is singular
    kw = dict(precond=bilu.GENERAL_ISAI, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **kw)
    assert_same(g, run_oracle(pat, A, b, **kw), "mixed general isai zero pivot")
    assert g[2][3] == bilu.ZERO_PIVOT and not g[0][3].any()


def test_mixed_rejects_unsupported():
    pat, A, b = synth.batch("random", 4, n=300, density=0.09)   # ~8400 nnz: beyond the persistent kernels' tables
    with pytest.raises(bilu.BiluError):
        run_gpu(pat, A, b, precision=bilu.PRECISION_FP32_IR, tol=1e-10, tol_mode=1)
    pat, A, b = synth.batch("random", 4, n=144)
    with pytest.raises(bilu.BiluError):   # the general ISAI has its own generation launch
        run_gpu(pat, A, b, precond=bilu.GENERAL_ISAI, precision=bilu.PRECISION_FP32_IR, merge=bilu.MERGED)


def test_mixed_full_batch_properties():
    print("case mixed full batch 65536", flush=True)
    B = 65536
    pat, A, b = synth.batch("random", B, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    x = torch.empty_like(db)
    kw = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
    it, cv, ir, tr, rf, fs = bilu.solve_ex(P, dA, db, x, bilu.opts(precision=bilu.PRECISION_FP32_IR, **kw))
    x64 = torch.empty_like(db)
    bilu.solve(P, dA, db, x64, bilu.opts(true_res_restart=1, **kw))
    res = torch.empty(B, dtype=torch.float64, device=dev())
    bilu.true_residual(P, dA, db, x, res)      # independent FP64 kernel
    torch.cuda.synchronize()
    bn = np.linalg.norm(b, axis=1)
    rel = res.cpu().numpy() / bn
    assert (cv.cpu().numpy() == 1).all()
    assert rel.max() <= 1e-10
    assert np.allclose(tr.cpu().numpy() / bn, rel, rtol=1e-3, atol=1e-12)
    err = (torch.linalg.norm(x - x64, dim=1) / torch.linalg.norm(x64, dim=1)).cpu().numpy()
    print(f"mixed vs FP64 solution: max rel error {err.max():.3e}, refine steps "
          f"{rf.cpu().numpy().mean():.2f} mean / {rf.cpu().numpy().max()} max", flush=True)
    assert err.max() <= 1e-6
    # spot check against the oracle
    idx = np.random.default_rng(0).choice(B, 64, replace=False)
    o = run_oracle(pat, A[idx], b[idx], precision=1, **kw)
    assert np.array_equal(x.cpu().numpy()[idx], o[0])


@pytest.mark.parametrize("kw", [
    dict(tol=1e-11, tol_mode=bilu.TOL_ABSOLUTE),                     # the reference's ATOL semantics
    dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, max_iter=2),         # inner solves cut short
    dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, max_refine=1),       # not converged after one step
    dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, inner_tol=0.5),      # loose inner solves, many steps
], ids=["absolute", "max_iter2", "max_refine1", "inner_tol_loose"])
@pytest.mark.parametrize("precond", [bilu.ILU0, bilu.ILU0_ISAI])
def test_mixed_options_bitwise(kw, precond):
    print(f"case mixed options {kw} precond={precond}", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    full = dict(kw, precond=precond, precision=bilu.PRECISION_FP32_IR)
    g = run_gpu(pat, A, b, **full)
    o = run_oracle(pat, A, b, **full)
    assert_same(g, o, f"mixed options {kw}")
    if kw.get("max_refine") == 1:
        assert (o[2] == bilu.NOT_CONVERGED).all() and (o[5] == 1).all()


@pytest.mark.parametrize("kind,B,n", [("random", 128, 144), ("stencil5", 64, 100), ("random", 64, 32),
                                      ("random", 64, 90)])
@pytest.mark.parametrize("precision,precond", [(bilu.PRECISION_FP64, bilu.PARILU0),
                                               (bilu.PRECISION_FP32_IR, bilu.PARILU0),
                                               (bilu.PRECISION_FP32_IR, bilu.PARILU0_ISAI)])
@pytest.mark.parametrize("stop", [0.0, 1e-8])
def test_inplace_parilu_bitwise(kind, B, n, precision, precond, stop):
    """In-place ParILU(0) sweeps (rounds reading what the earlier rounds of
    the sweep wrote) through the persistent factor kernel, fixed count or
    stopped on the residual, bit for bit against the oracle's in-place sweep."""
    print(f"case inplace parilu {kind} n={n} precision={precision} precond={precond} stop={stop}", flush=True)
    pat, A, b = synth.batch(kind, B, n=n)
    kw = dict(precond=precond, parilu_sweeps=40 if stop > 0 else 3, parilu_stop_tol=stop, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=precision, parilu_order=bilu.PARILU_INPLACE)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"inplace parilu {kind} n={n}")
    assert (o[2] == 1).all()
    if stop > 0:
        assert (o[6] >= 1).all()


def test_inplace_parilu_warm_bitwise():
    """In-place sweeps from a caller's warm start (perturbed exact factors)."""
    print("case inplace parilu warm", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    F = np.stack([O.ilu0(pat, A[s])[0] for s in range(A.shape[0])])
    init = F * (1.0 + 1e-3 * np.random.default_rng(11).standard_normal(F.shape))
    kw = dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR, parilu_order=bilu.PARILU_INPLACE)
    g = run_gpu(pat, A, b, F_init=init, **kw)
    assert_same(g, run_oracle(pat, A, b, init=init, **kw), "inplace parilu warm")
    # fewer in-place sweeps than Jacobi ones from the same start
    kj = dict(kw, parilu_order=bilu.PARILU_JACOBI)
    gj = run_gpu(pat, A, b, F_init=init, **kj)
    assert g[6].mean() < gj[6].mean(), (g[6].mean(), gj[6].mean())


@pytest.mark.parametrize("inplace", [False, True])
def test_fp64_isai_split_bitwise(inplace):
    """FP64 ISAI(L, U) after residual-stopped ParILU(0) (Jacobi or in place):
    the persistent factor + ISAI launch, then the persistent FP64 solve."""
    print(f"case fp64 isai split inplace={inplace}", flush=True)
    pat, A, b = synth.batch("random", 128, n=144)
    kw = dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE,
              parilu_order=bilu.PARILU_INPLACE if inplace else bilu.PARILU_JACOBI)
    g = run_gpu(pat, A, b, **kw)
    o = run_oracle(pat, A, b, **kw)
    assert_same(g, o, f"fp64 isai split inplace={inplace}")
    assert (o[2] == 1).all() and (o[6] > 0).all()


MERGE_CASES = [
    ("ilu0", dict(precond=bilu.ILU0)),
    ("ilu0_isai", dict(precond=bilu.ILU0_ISAI)),
    ("parilu0_jacobi_5", dict(precond=bilu.PARILU0, parilu_sweeps=5)),
    ("parilu0_inplace_stop", dict(precond=bilu.PARILU0, parilu_sweeps=40, parilu_stop_tol=1e-8,
                                  parilu_order=bilu.PARILU_INPLACE)),
    ("parilu0_isai_jacobi_stop", dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8)),
    ("parilu0_isai_inplace_stop", dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8,
                                       parilu_order=bilu.PARILU_INPLACE)),
]


@pytest.mark.parametrize("kind,B,n", [("random", 192, 144), ("stencil5", 64, 100), ("random", 64, 32),
                                      ("random", 48, 300)])
@pytest.mark.parametrize("name,kw", MERGE_CASES, ids=[c[0] for c in MERGE_CASES])
def test_mixed_merged_bitwise(kind, B, n, name, kw):
    """the merged kernel (each system factored in the wave's LDS, then solved)
    against the oracle and the segregated launches: every output bit for bit"""
    print(f"case mixed merged {kind} B={B} n={n} {name}", flush=True)
    dens = {} if kind != "random" or n == 144 else dict(density=13.0 / (n - 1))
    pat, A, b = synth.batch(kind, B, n=n, **dens)
    base = dict(kw, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    res = {}
    for merge in (bilu.MERGED, bilu.SEGREGATED):
        x = torch.empty_like(db)
        fs = torch.full((B,), -7, dtype=torch.int32, device=dev())
        out = bilu.solve_ex(P, dA, db, x, bilu.opts(**base, merge=merge), factor_status=fs)
        torch.cuda.synchronize()
        res[merge] = (x.cpu().numpy(),) + tuple(t.cpu().numpy() for t in out)
    assert_same(res[bilu.MERGED], res[bilu.SEGREGATED], f"merged vs segregated {name}")
    assert_same(res[bilu.MERGED], run_oracle(pat, A, b, **base), f"merged vs oracle {name}")
    assert (res[bilu.MERGED][2] == 1).all()


def test_mixed_merged_warm_and_zero_pivot():
    """merged: the caller's ParILU guess (F_init) and a zero pivot, as in the segregated launches"""
    print("case mixed merged warm start + zero pivot", flush=True)
    pat, A, b = synth.batch("random", 96, n=144)
    F = np.stack([O.ilu0(pat, A[s])[0] for s in range(A.shape[0])])
    init = F * (1.0 + 1e-3 * np.random.default_rng(11).standard_normal(F.shape))
    for pc in (bilu.PARILU0, bilu.PARILU0_ISAI):
        kw = dict(precond=pc, parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10, tol_mode=bilu.TOL_RELATIVE,
                  precision=bilu.PRECISION_FP32_IR)
        g = run_gpu(pat, A, b, F_init=init, merge=bilu.MERGED, **kw)
        assert_same(g, run_gpu(pat, A, b, F_init=init, merge=bilu.SEGREGATED, **kw), f"merged warm {pc}")
        assert_same(g, run_oracle(pat, A, b, init=init, **kw), f"merged warm vs oracle {pc}")
    A = A.copy()
    A[5][pat.diag[3]] = 0.0
    for pc in (bilu.ILU0, bilu.ILU0_ISAI):
        kw = dict(precond=pc, tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
        g = run_gpu(pat, A, b, merge=bilu.MERGED, **kw)
        assert_same(g, run_gpu(pat, A, b, merge=bilu.SEGREGATED, **kw), f"merged zero pivot {pc}")
        assert_same(g, run_oracle(pat, A, b, **kw), f"merged zero pivot vs oracle {pc}")
        assert g[2][5] == bilu.ZERO_PIVOT and (np.delete(g[2], 5) == 1).all()
