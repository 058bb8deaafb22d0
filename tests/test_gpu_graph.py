"""Solves captured into a hipGraph after bilu_pattern_reserve: the capture
records only kernel launches (no allocation, no device wait, no memset), and
every replay -- with eager work and synchronisations between replays --
reproduces the eager solve bit for bit."""
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
    "fp64_merged_ilu0": dict(precond=bilu.ILU0, true_res_restart=1),
    "fp64_segregated_ilu0": dict(precond=bilu.ILU0, merge=bilu.SEGREGATED),
    "mixed_ilu0_isai": dict(precond=bilu.ILU0_ISAI, precision=bilu.PRECISION_FP32_IR),
    "mixed_parilu_inplace_stop": dict(precond=bilu.PARILU0_ISAI, precision=bilu.PRECISION_FP32_IR,
                                      parilu_sweeps=40, parilu_stop_tol=1e-8,
                                      parilu_order=bilu.PARILU_INPLACE),
}


@pytest.mark.parametrize("case", list(CASES))
def test_solve_graph_replays_match_eager(case):
    print(f"case graph {case}", flush=True)
    pat, A, b = synth.batch("random", 512, n=144)
    P = bilu.Pattern.from_synth(pat)
    P.reserve(A.shape[0])
    dA = torch.from_numpy(A).to(dev())
    db = torch.from_numpy(b).to(dev())
    o = bilu.opts(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, **CASES[case])
    B = A.shape[0]

    par = CASES[case]["precond"] in (bilu.PARILU0, bilu.PARILU0_ISAI)

    def outputs():
        # the factor status only for ParILU (asking for it routes a plain solve through the split launches)
        return [torch.empty(B, dtype=torch.float32, device=dev()), torch.empty(B, dtype=torch.int32, device=dev()),
                torch.empty(B, dtype=torch.float64, device=dev()), torch.empty(B, dtype=torch.float64, device=dev()),
                torch.empty(B, dtype=torch.int32, device=dev())] + \
               ([torch.empty(B, dtype=torch.int32, device=dev())] if par else [False])

    x0 = torch.empty_like(db)
    out0 = outputs()
    bilu.solve_ex(P, dA, db, x0, o, None, *out0)
    torch.cuda.synchronize()
    assert (out0[1].cpu().numpy() == 1).all()

    x1 = torch.empty_like(db)
    out1 = outputs()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        bilu.solve_ex(P, dA, db, x1, o, None, *out1, stream=s)
    for rep in range(3):
        x1.fill_(np.nan)
        for t in out1:
            if t is not False:
                t.fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(x1, x0), f"replay {rep}: x differs"
        for k, (u, v) in enumerate(zip(out1, out0)):
            assert u is False or torch.equal(u, v), f"replay {rep}: output {k} differs"
        # eager work between replays
        bilu.solve_ex(P, dA, db, torch.empty_like(db), o, None, *outputs())
        torch.cuda.synchronize()
