"""The committed golden fixtures (tests/golden/*.npz, written by
make_golden.py from the oracle): the oracle must keep reproducing them bit
for bit (CPU), and the library's FP64 factorization, ParILU(0) and ISAI
entry points must match them (GPU).  Inputs are regenerated from the
counter-based generator, so only outputs are stored."""
import os

import numpy as np
import pytest

import oracle as O
import synth

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"stencil5_b64": ("stencil5", 64, 100), "random_n144_b8": ("random", 8, 144),
         "random_n32_b8": ("random", 8, 32), "random_n300_b4": ("random", 4, 300)}


def load(name):
    kind, B, n = CASES[name]
    pat, A, b = synth.batch(kind, B, n=n)
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    assert np.array_equal(g["row_ptrs"], pat.row_ptrs) and np.array_equal(g["col_idx"], pat.col_idx)
    return pat, A, b, g


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    pat, A, b, g = load(name)
    for tag, o in (("ref", O.reference_opts()), ("w64", O.opts()),
                   ("rel", O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1))):
        x, it, cv, ir, tr = O.batch_solve(pat, A, b, O.ILU0, o=o)
        for key, v in (("x", x), ("iter", it), ("conv", cv), ("iter_res", ir), ("true_res", tr)):
            assert np.array_equal(v, g[f"{key}_{tag}"]), f"{name} {key}_{tag}"
    k = g["ilu0"].shape[0]
    assert np.array_equal(np.stack([O.ilu0(pat, A[s])[0] for s in range(k)]), g["ilu0"])
    assert np.array_equal(np.stack([O.parilu0(pat, A[s], 5) for s in range(k)]), g["parilu5"])
    assert np.array_equal(O.isai(pat, g["ilu0"][0], O.ISAI_LOWER)[0], g["isai_lower"])
    assert np.array_equal(O.isai(pat, g["ilu0"][0], O.ISAI_UPPER)[0], g["isai_upper"])
    if "isai_general" in g.files:
        assert np.array_equal(O.isai(pat, A[0], O.ISAI_GENERAL)[0], g["isai_general"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_library_matches_golden(name):
    torch = pytest.importorskip("torch")
    import bilu
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    print(f"case golden {name}", flush=True)
    pat, A, b, g = load(name)
    k = g["ilu0"].shape[0]
    P = bilu.Pattern.from_synth(pat)
    dA = torch.from_numpy(np.ascontiguousarray(A[:k])).to("cuda")
    F = torch.empty_like(dA)
    bilu.ilu0(P, dA, F)
    F5 = torch.empty_like(dA)
    bilu.parilu0(P, dA, F5, 5)
    torch.cuda.synchronize()
    assert np.array_equal(F.cpu().numpy(), g["ilu0"]), f"{name}: ILU(0) factors"
    assert np.array_equal(F5.cpu().numpy(), g["parilu5"]), f"{name}: 5 Jacobi ParILU(0) sweeps"
    for typ, key, src in ((bilu.ISAI_LOWER, "isai_lower", g["ilu0"][0]), (bilu.ISAI_UPPER, "isai_upper", g["ilu0"][0]),
                          (bilu.ISAI_GENERAL, "isai_general", A[0])):
        if key not in g.files:
            continue
        vals = torch.from_numpy(np.ascontiguousarray(src[None, :])).to("cuda")
        out = torch.zeros_like(vals)
        bilu.isai(P, vals, typ, out)
        torch.cuda.synchronize()
        mask = np.zeros(pat.nnz, bool)
        rows = np.repeat(np.arange(pat.n), np.diff(pat.row_ptrs))
        mask[:] = True if typ == bilu.ISAI_GENERAL else (rows > pat.col_idx if typ == bilu.ISAI_LOWER
                                                          else rows <= pat.col_idx)
        assert np.array_equal(out.cpu().numpy()[0][mask], g[key][mask]), f"{name}: {key}"
