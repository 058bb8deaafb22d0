"""Missing-diagonal insertion (reference factorization.cu:399-467,
Count_Missing_Diagonal_Elements + Add_Missing_Diagonal_Elements, used by
ILU_0_Factorization_Gpu and ParILU_0_Factorization_Gpu).

CPU: the oracle's restatement is pinned by brute force (the sorted union of
each row's columns with i) and by the dense identity of the expanded values;
the library's host pattern function equals the oracle bit for bit and
rejects malformed patterns.  GPU: bilu_expand_values equals the oracle's
gather, and the drop-in factorizations / ILU solvers accept patterns
without a full diagonal with the oracle's results on the augmented pattern,
while the merged / segregated / general ISAI solvers reject them."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import bilu
import oracle as O
import synth


def brute_force(n, rp, ci):
    rows = [sorted(set(ci[rp[i]:rp[i + 1]].tolist()) | {i}) for i in range(n)]
    rpo = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return rpo, np.array([c for r in rows for c in r], np.int32)


def random_csr(n, rng, density, drop_diag):
    """random sorted CSR; the diagonal of row i kept with probability 1 - drop_diag"""
    rows = []
    for i in range(n):
        cols = set(np.nonzero(rng.random(n) < density)[0].tolist())
        cols.discard(i)
        if rng.random() >= drop_diag:
            cols.add(i)
        rows.append(sorted(cols))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return rp, np.array([c for r in rows for c in r], np.int32)


def edge_patterns():
    """rows without entries, all columns left / right of i, diagonal missing
    in the first / last row, a single row, rows of more than 64 entries"""
    yield 1, np.array([0, 0], np.int32), np.array([], np.int32)
    yield 1, np.array([0, 1], np.int32), np.array([0], np.int32)
    yield 4, np.array([0, 2, 2, 4, 5], np.int32), np.array([1, 3, 0, 1, 0], np.int32)
    yield 3, np.array([0, 2, 4, 6], np.int32), np.array([1, 2, 0, 2, 0, 1], np.int32)
    n = 100
    rows = [[c for c in range(n) if c != i and (c % 3 == 0 or i % 10 == 0)] for i in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    yield n, rp, np.array([c for r in rows for c in r], np.int32)


def cases():
    rng = np.random.default_rng(11)
    yield from edge_patterns()
    for n, dens, drop in [(10, 0.3, 0.5), (50, 0.1, 0.3), (144, 0.05, 0.1), (200, 0.4, 0.9), (64, 0.0, 1.0)]:
        yield (n, *random_csr(n, rng, dens, drop))


@pytest.mark.parametrize("case", list(range(10)))
def test_oracle_insertion_matches_brute_force(case):
    n, rp, ci = list(cases())[case]
    rpo, cio, h = O.add_diagonal(n, rp, ci)
    want_rp, want_ci = brute_force(n, rp, ci)
    assert np.array_equal(rpo, want_rp) and np.array_equal(cio, want_ci)
    # holders: inserted entries are exactly the diagonals that were missing
    for i in range(n):
        for k in range(rpo[i], rpo[i + 1]):
            if h[k] < 0:
                assert h[k] == -i - 1 and cio[k] == i and i not in ci[rp[i]:rp[i + 1]]
            else:
                assert ci[h[k]] == cio[k] and rp[i] <= h[k] < rp[i + 1]


@pytest.mark.parametrize("case", list(range(10)))
def test_oracle_expanded_values_keep_the_matrix(case):
    n, rp, ci = list(cases())[case]
    rng = np.random.default_rng(case)
    vals = rng.standard_normal((3, len(ci)))
    rpo, cio, h = O.add_diagonal(n, rp, ci)
    out = O.expand_values(h, vals)
    P0, P1 = synth.Pattern(n, rp, ci), synth.Pattern(n, rpo, cio)
    for s in range(3):
        assert np.array_equal(P1.dense(out[s]), P0.dense(vals[s]))
    assert np.all(out[:, h < 0] == 0.0)


def test_oracle_round_trip_of_removed_diagonals():
    for kind, n in [("stencil5", 100), ("random", 144)]:
        pat, A, b = synth.batch(kind, 4, n=n)
        rp, ci, Ak = remove_diagonals(pat, A, range(3, pat.n, 11))
        rpo, cio, h = O.add_diagonal(pat.n, rp, ci)
        assert np.array_equal(rpo, pat.row_ptrs) and np.array_equal(cio, pat.col_idx)
        Aa = O.expand_values(h, Ak)
        gone = np.isin(np.arange(pat.nnz), pat.diag[list(range(3, pat.n, 11))])
        assert np.array_equal(Aa[:, ~gone], A[:, ~gone]) and np.all(Aa[:, gone] == 0.0)


@pytest.mark.parametrize("case", list(range(10)))
def test_library_insertion_matches_oracle(case):
    n, rp, ci = list(cases())[case]
    rpo, cio, src = bilu.add_diagonal(n, rp, ci)
    want_rp, want_ci, h = O.add_diagonal(n, rp, ci)
    assert np.array_equal(rpo, want_rp) and np.array_equal(cio, want_ci)
    assert np.array_equal(src, np.where(h < 0, -1, h))


def test_library_full_diagonal_is_unchanged():
    pat, _, _ = synth.batch("random", 1, n=144)
    rpo, cio, src = bilu.add_diagonal(pat.n, pat.row_ptrs, pat.col_idx)
    assert np.array_equal(rpo, pat.row_ptrs) and np.array_equal(cio, pat.col_idx)
    assert np.array_equal(src, np.arange(pat.nnz))


@pytest.mark.parametrize("rp,ci,code", [
    ([0, 2, 3], [1, 0, 0], bilu.ERR_INVALID_PATTERN),   # unsorted row
    ([0, 1, 3], [0, 0, 2], bilu.ERR_INVALID_PATTERN),   # column out of range
    ([0, 2, 1], [0, 1], bilu.ERR_INVALID_PATTERN),      # row_ptrs decreasing / rp[n] != nnz
    ([1, 2, 2], [0, 1], bilu.ERR_INVALID_PATTERN),      # rp[0] != 0
])
def test_library_rejects_malformed_patterns(rp, ci, code):
    rp, ci = np.array(rp, np.int32), np.array(ci, np.int32)
    P = C.POINTER(C.c_int)
    miss = C.c_int()
    rc = bilu.lib().bilu_pattern_add_diagonal(2, len(ci), rp.ctypes.data_as(P), ci.ctypes.data_as(P), C.byref(miss),
                                              None, None, None)
    assert rc == code
    assert bilu.lib().bilu_pattern_add_diagonal(0, 0, rp.ctypes.data_as(P), ci.ctypes.data_as(P), C.byref(miss),
                                                None, None, None) == bilu.ERR_INVALID_ARG


def remove_diagonals(pat, A, rows_to_drop):
    rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
    keep = ~((rows == pat.col_idx) & np.isin(rows, list(rows_to_drop)))
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=pat.n))]).astype(np.int32)
    return rp, pat.col_idx[keep].astype(np.int32), np.ascontiguousarray(A[:, keep])


@pytest.mark.gpu
def test_expand_values_matches_oracle_full_batch():
    import torch
    print("case expand_values n=144 B=65536", flush=True)
    pat, A, _ = synth.batch("random", 65536, n=144)
    rp, ci, Ak = remove_diagonals(pat, A, range(0, pat.n, 5))
    rpo, cio, src = bilu.add_diagonal(pat.n, rp, ci)
    dev = torch.device("cuda:0")
    d_src = torch.from_numpy(src).to(dev)
    d_in = torch.from_numpy(Ak).to(dev)
    d_out = torch.full((65536, len(cio)), float("nan"), dtype=torch.float64, device=dev)
    bilu.expand_values(d_src, d_in, d_out)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    want = np.where(src < 0, 0.0, Ak[:, np.maximum(src, 0)])  # the oracle's gather, vectorised
    assert np.array_equal(got, want)
    small = O.expand_values(np.where(src < 0, -1, src), Ak[:64])  # and the C oracle on a slice
    assert np.array_equal(got[:64], small)


HERE = os.path.dirname(os.path.abspath(bilu.__file__))
DRIVER = os.path.join(HERE, "bicgstab_driver")


def _run_driver(tmp_path, variant, rp, ci, Ak, b, n):
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, len(ci), Ak.shape[0]], np.int32).tofile(f)
        rp.tofile(f)
        ci.tofile(f)
        Ak.tofile(f)
        np.ascontiguousarray(b).tofile(f)
    r = subprocess.run([DRIVER, variant, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    return r, fout


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["ilu0", "parilu0"])
def test_dropin_factorizations_insert_diagonals(tmp_path, variant):
    from test_dropin import read_factors
    print(f"case missing-diagonal factor {variant}", flush=True)
    pat, A, b = synth.batch("stencil5", 32, n=100)
    drop = range(3, pat.n, 11)
    rp, ci, Ak = remove_diagonals(pat, A, drop)
    r, fout = _run_driver(tmp_path, variant, rp, ci, Ak, b, pat.n)
    assert r.returncode == 0, r.stderr
    lrp, lci, urp, uci, Lv, Uv = read_factors(fout, 32, pat.n)
    Aa = O.expand_values(O.add_diagonal(pat.n, rp, ci)[2], Ak)  # values on the full stencil pattern
    for s in range(32):
        F = O.ilu0(pat, Aa[s])[0] if variant == "ilu0" else O.parilu0(pat, Aa[s], 5)
        Lw, Uw = [], []
        for i in range(pat.n):
            Lw += list(F[pat.row_ptrs[i]:pat.diag[i]]) + [1.0]
            Uw += list(F[pat.diag[i]:pat.row_ptrs[i + 1]])
        assert np.array_equal(Lv[s], Lw, equal_nan=True) and np.array_equal(Uv[s], Uw, equal_nan=True), s
    assert len(lci) + len(uci) == pat.nnz + pat.n


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n", [("stencil5", 100), ("random", 144)])
@pytest.mark.parametrize("variant", ["ilu", "parilu", "ilu_isai"])
def test_dropin_ilu_solvers_insert_diagonals(tmp_path, kind, n, variant):
    from test_dropin import dropin_oracle, read_result
    print(f"case missing-diagonal solve {variant} {kind}", flush=True)
    pat, A, b = synth.batch(kind, 48, n=n)
    rp, ci, Ak = remove_diagonals(pat, A, range(3, pat.n, 11))
    r, fout = _run_driver(tmp_path, variant, rp, ci, Ak, b, pat.n)
    assert r.returncode == 0, r.stderr
    x, it, cv, ir, tr, _ = read_result(fout, 48, pat.n)
    Aa = O.expand_values(O.add_diagonal(pat.n, rp, ci)[2], Ak)
    o = dropin_oracle(pat, Aa, b, variant)
    for name, u, v in zip(["x", "iters", "conv", "iter_res", "true_res"], (x, it, cv, ir, tr), o):
        assert np.array_equal(u, v, equal_nan=True), f"{variant}: {name} differs"
    if kind == "stencil5" and variant != "parilu":
        assert np.all(cv == 1)  # zero diagonals filled by the factorization: converges


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["merged", "segregated", "general_isai"])
def test_dropin_solvers_reject_missing_diagonal(tmp_path, variant):
    print(f"case missing-diagonal reject {variant}", flush=True)
    pat, A, b = synth.batch("stencil5", 8, n=100)
    rp, ci, Ak = remove_diagonals(pat, A, [5])
    r, _ = _run_driver(tmp_path, variant, rp, ci, Ak, b, pat.n)
    assert r.returncode == 1 and "missing diagonal" in r.stderr, r.stderr
