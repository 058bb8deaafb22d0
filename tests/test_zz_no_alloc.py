"""bilu_solve_ex_noalloc (the library's fixed scratch array instead of the
handle workspace) against bilu_solve_ex through the C ABI: every output bit
for bit, two calls in a row, the second with a smaller batch, on the paths
that pass factors or ISAI values through the scratch (mixed precision,
split FP64 ParILU, segregated FP64) and ones that need none (merged FP64,
the merged mixed-precision kernel).
The first call's batch spans several scratch chunks."""
import numpy as np
import pytest

import bilu
import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


CASES = {
    "mixed_ilu0_isai": dict(precond=bilu.ILU0_ISAI, precision=bilu.PRECISION_FP32_IR),
    "mixed_parilu_inplace_stop": dict(precond=bilu.PARILU0_ISAI, precision=bilu.PRECISION_FP32_IR,
                                      parilu_sweeps=40, parilu_stop_tol=1e-8, parilu_order=bilu.PARILU_INPLACE),
    "fp64_split_parilu": dict(precond=bilu.PARILU0, parilu_sweeps=40, parilu_stop_tol=1e-8),
    "fp64_segregated_ilu0": dict(precond=bilu.ILU0, merge=bilu.SEGREGATED),
    "fp64_merged_ilu0": dict(precond=bilu.ILU0, true_res_restart=1),
    "mixed_merged_kernel": dict(precond=bilu.ILU0_ISAI, precision=bilu.PRECISION_FP32_IR, merge=bilu.MERGED),
}

# 36000 systems at nnz 2082 need three chunks of the 256 MiB scratch
SIZES = (36000, 700)


def run(fn, P, dA, db, o, par):
    B = dA.shape[0]
    x = torch.empty_like(db)
    outs = [torch.empty(B, dtype=torch.float32, device=dev()), torch.empty(B, dtype=torch.int32, device=dev()),
            torch.empty(B, dtype=torch.float64, device=dev()), torch.empty(B, dtype=torch.float64, device=dev()),
            torch.empty(B, dtype=torch.int32, device=dev()),
            torch.empty(B, dtype=torch.int32, device=dev()) if par else False]
    fn(P, dA, db, x, o, None, *outs)
    torch.cuda.synchronize()
    return x, outs


@pytest.mark.parametrize("case", list(CASES))
def test_noalloc_matches_solve_ex(case):
    kw = CASES[case]
    par = kw["precond"] in (bilu.PARILU0, bilu.PARILU0_ISAI)
    o = bilu.opts(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, **kw)
    pat, A, b = synth.batch("random", SIZES[0], first=5 << 20, n=144)
    P = bilu.Pattern.from_synth(pat)
    for B in SIZES:
        print(f"case noalloc {case} batch {B}", flush=True)
        dA = torch.from_numpy(A[:B]).to(dev())
        db = torch.from_numpy(b[:B]).to(dev())
        x1, o1 = run(bilu.solve_ex_noalloc, P, dA, db, o, par)
        x0, o0 = run(bilu.solve_ex, P, dA, db, o, par)
        assert torch.equal(x1, x0), f"{case} batch {B}: x differs"
        for k, (u, v) in enumerate(zip(o1, o0)):
            assert u is False or torch.equal(u, v), f"{case} batch {B}: output {k} differs"
        assert (o0[1].cpu().numpy() == 1).all()


def test_noalloc_rejects_power_isai():
    pat, A, b = synth.batch("random", 4, n=32)
    P = bilu.Pattern.from_synth(pat)
    dA, db = torch.from_numpy(A).to(dev()), torch.from_numpy(b).to(dev())
    with pytest.raises(bilu.BiluError):
        bilu.solve_ex_noalloc(P, dA, db, torch.empty_like(db), bilu.opts(precond=bilu.ILU0_ISAI, isai_power=2))
