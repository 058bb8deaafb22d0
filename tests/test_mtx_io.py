"""Matrix Market batch input / output of the drop-in interface
(ReadWriteData: FillSubdir, ReadData, Print_ans, PrintSolverResults, after
the reference's ReadWriteData.cpp) through the bicgstab_driver program.

CPU: a tree of problems is read back (convert) exactly as written -- general
and symmetric coordinate files, shuffled entries, comments, integer values,
the scaled file names, Problem_Size copies -- and malformed trees are
rejected.  GPU: solving from a tree writes one answer file per problem whose
values equal the solver's x bit for bit, and a results file with the
convergence flag and iteration count of each problem."""
import os
import subprocess

import numpy as np
import pytest

import bilu
import synth

HERE = os.path.dirname(os.path.abspath(bilu.__file__))
DRIVER = os.path.join(HERE, "bicgstab_driver")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="bicgstab_driver not built")


def write_coo(path, pat, vals, symmetric=False, seed=None, integer=False):
    rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
    cols = pat.col_idx
    keep = rows >= cols if symmetric else np.ones(pat.nnz, bool)
    r, c, v = rows[keep], cols[keep], vals[keep]
    order = np.random.default_rng(seed).permutation(len(r)) if seed is not None else np.arange(len(r))
    with open(path, "w") as f:
        kind = "integer" if integer else "real"
        f.write(f"%%MatrixMarket matrix coordinate {kind} {'symmetric' if symmetric else 'general'}\n")
        f.write("% written by the Matrix Market input test\n%\n")
        f.write(f"{pat.n} {pat.n} {len(r)}\n")
        for k in order:
            val = str(int(v[k])) if integer else repr(float(v[k]))
            f.write(f"{r[k] + 1} {c[k] + 1} {val}\n")


def write_array(path, vec):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix array real general\n{len(vec)} 1\n")
        for x in vec:
            f.write(repr(float(x)) + "\n")


def read_array(path):
    lines = [ln for ln in open(path).read().split("\n") if ln and not ln.startswith("%")]
    m, one = map(int, lines[0].split())
    assert one == 1 and len(lines) == m + 1
    return np.array([float(x) for x in lines[1:]])


def symmetric_values(pat, vals):
    D = pat.dense(vals)
    S = 0.5 * (D + D.T)
    rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
    return S[rows, pat.col_idx]


def make_tree(root, pat, A, b, scaled=False, symmetric_first=False, integer_second=False):
    """one subdirectory per system, named so that sorted order is system order"""
    names = []
    A = A.copy()
    for s in range(A.shape[0]):
        d = root / f"p{s:03d}"
        d.mkdir(parents=True)
        sym = symmetric_first and s == 0
        integer = integer_second and s == 1
        if sym:
            A[s] = symmetric_values(pat, A[s])
        if integer:
            A[s] = np.round(A[s] * 8)
        write_coo(d / ("A_scaled.mtx" if scaled else "A.mtx"), pat, A[s], symmetric=sym, seed=s, integer=integer)
        write_array(d / ("b_scaled.mtx" if scaled else "b.mtx"), b[s])
        names.append(d)
    return A, names


def read_bin(path):
    raw = open(path, "rb").read()
    n, nnz, pages = np.frombuffer(raw, np.int32, 3)
    o = 12
    rp = np.frombuffer(raw, np.int32, n + 1, o); o += rp.nbytes
    ci = np.frombuffer(raw, np.int32, nnz, o); o += ci.nbytes
    v = np.frombuffer(raw, np.float64, pages * nnz, o).reshape(pages, nnz); o += v.nbytes
    b = np.frombuffer(raw, np.float64, pages * n, o).reshape(pages, n); o += b.nbytes
    assert o == len(raw)
    return rp, ci, v, b


def convert(tree, problem_size, scaled, out):
    return subprocess.run([DRIVER, "convert", "--mtx", str(tree), str(problem_size), "1" if scaled else "0", str(out)],
                          capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("kind,n,scaled", [("stencil5", 100, False), ("random", 144, True)])
def test_tree_reads_back_exactly(tmp_path, kind, n, scaled):
    pat, A, b = synth.batch(kind, 5, n=n)
    A, _ = make_tree(tmp_path / "cat", pat, A, b, scaled=scaled, symmetric_first=(kind == "stencil5"),
                     integer_second=True)
    r = convert(tmp_path / "cat", 3, scaled, tmp_path / "out.bin")
    assert r.returncode == 0, r.stderr
    rp, ci, v, bb = read_bin(tmp_path / "out.bin")
    assert np.array_equal(rp, pat.row_ptrs) and np.array_equal(ci, pat.col_idx)
    assert v.shape == (15, pat.nnz)
    for p in range(15):  # page p holds problem p % 5
        assert np.array_equal(v[p], A[p % 5]), p
        assert np.array_equal(bb[p], b[p % 5]), p


def test_sorted_subdirectories_and_trailing_slash(tmp_path):
    pat, A, b = synth.batch("random", 3, n=60)
    root = tmp_path / "cat"
    for s, name in enumerate(["zeta", "alpha", "mid"]):  # read order is alphabetical
        d = root / name
        d.mkdir(parents=True)
        write_coo(d / "A.mtx", pat, A[s], seed=7)
        write_array(d / "b.mtx", b[s])
    r = subprocess.run([DRIVER, "convert", "--mtx", str(root) + "/", "1", "0", str(tmp_path / "o.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    _, _, v, bb = read_bin(tmp_path / "o.bin")
    assert np.array_equal(v, A[[1, 2, 0]]) and np.array_equal(bb, b[[1, 2, 0]])


def _bad_tree(tmp_path, what):
    pat, A, b = synth.batch("random", 2, n=40)
    root = tmp_path / "cat"
    make_tree(root, pat, A, b)
    p1 = root / "p001"
    if what == "pattern":
        other = synth.random_pattern(40, 0.1, seed=99)
        write_coo(p1 / "A.mtx", other, np.ones(other.nnz))
    elif what == "missing_b":
        os.remove(p1 / "b.mtx")
    elif what == "b_length":
        write_array(p1 / "b.mtx", b[1][:-1])
    elif what == "nonsquare":
        txt = open(p1 / "A.mtx").read().replace("\n40 40 ", "\n40 41 ", 1)
        open(p1 / "A.mtx", "w").write(txt)
    elif what == "range":
        with open(p1 / "A.mtx") as f:
            lines = f.read().split("\n")
        lines[4] = "41 1 1.0"
        open(p1 / "A.mtx", "w").write("\n".join(lines))
    elif what == "duplicate":
        with open(p1 / "A.mtx") as f:
            lines = f.read().rstrip("\n").split("\n")
        lines[5] = lines[4]
        open(p1 / "A.mtx", "w").write("\n".join(lines) + "\n")
    elif what == "short":
        with open(p1 / "A.mtx") as f:
            lines = f.read().rstrip("\n").split("\n")
        open(p1 / "A.mtx", "w").write("\n".join(lines[:-3]) + "\n")
    elif what == "banner":
        txt = open(p1 / "A.mtx").read().replace("%%MatrixMarket matrix coordinate real", "%%MatrixMarket matrix "
                                                "coordinate complex", 1)
        open(p1 / "A.mtx", "w").write(txt)
    return root


@pytest.mark.parametrize("what,message", [
    ("pattern", "sparsity pattern differs"), ("missing_b", "cannot open"), ("b_length", "n x 1 array"),
    ("nonsquare", "not square"), ("range", "index out of range"), ("duplicate", "duplicate entry"),
    ("short", "short file"), ("banner", "only real / integer")])
def test_malformed_trees_are_rejected(tmp_path, what, message):
    root = _bad_tree(tmp_path, what)
    r = convert(root, 1, False, tmp_path / "out.bin")
    assert r.returncode == 2, (what, r.stdout, r.stderr)
    assert "reading the batch failed" in r.stderr and message in r.stderr, r.stderr
    assert not os.path.exists(tmp_path / "out.bin")


def test_missing_category_directory(tmp_path):
    r = convert(tmp_path / "nope", 1, False, tmp_path / "out.bin")
    assert r.returncode == 2 and "could not open directory" in r.stderr


ANSWER_FILES = {"merged": "conv_ilu_merged_app1_bicgstab", "segregated": "conv_ilu_segregated_app1_bicgstab",
                "ilu": "conv_ILU_app_1_bicgstab", "parilu": "parILU_5_bicgstab", "ilu_isai": "ilu_isai_bicgstab",
                "parilu_isai": "ilu_isai_bicgstab", "general_isai": "general_isai_bicgstab"}


@pytest.mark.gpu
@pytest.mark.parametrize("variant,scaled", [("merged", False), ("segregated", True), ("ilu", False),
                                            ("parilu", False), ("ilu_isai", True), ("general_isai", False)])
def test_solve_from_tree_writes_answers(tmp_path, variant, scaled):
    from test_dropin import read_result
    print(f"case mtx tree {variant} scaled={scaled}", flush=True)
    pat, A, b = synth.batch("random", 6, n=144)
    A, dirs = make_tree(tmp_path / "cat", pat, A, b, scaled=scaled)
    work = tmp_path / "work"
    work.mkdir()
    r = subprocess.run([DRIVER, variant, "--mtx", str(tmp_path / "cat"), "2", "1" if scaled else "0",
                        str(work / "out.bin")], capture_output=True, text=True, timeout=120, cwd=work)
    assert r.returncode == 0, r.stderr
    x, it, cv, ir, tr, _ = read_result(work / "out.bin", 12, pat.n)
    # the batch read from the tree solves exactly like the same arrays given directly
    from test_dropin import dropin_oracle
    Ab, bb = np.concatenate([A, A]), np.concatenate([b, b])
    o = dropin_oracle(pat, Ab, bb, variant)
    assert np.array_equal(x, o[0]) and np.array_equal(it, o[1]) and np.array_equal(cv, o[2])
    # Problem_Size copies of a problem are the same system: identical answers, so one
    # answer file per problem directory loses nothing
    assert np.array_equal(x[6:], x[:6]) and np.array_equal(it[6:], it[:6])
    # one answer file per problem directory, the first pages of x, %0.17lg round trips
    name = ("x_scaled_gpu_" if scaled else "x_gpu_") + ANSWER_FILES[variant] + ".mtx"
    for s, d in enumerate(dirs):
        assert np.array_equal(read_array(d / name), x[s]), (d, name)
    # results: "flag   iterations" per problem directory
    lines = open(work / f"Results_{variant}_bicgstab_cat.txt").read().strip("\n").split("\n")
    assert len(lines) == len(dirs)
    for s, ln in enumerate(lines):
        flag, iters = ln.split()
        assert int(flag) == cv[s] and float(iters) == pytest.approx(float(it[s]), abs=5e-7)
