"""Uneven batches (several patterns, one interleaved request list) through
the hipGraph-captured plan: bitwise the same as solving each pattern's
systems on their own, unchanged by the autotuner's choice, stable over graph
replays, and the oracle's results for the 72-slot (n = 300) kernels, LU
and ISAI."""
import numpy as np
import pytest

import bilu
import oracle as O
import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NS = (32, 90, 144, 300)


def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def c3(B_each, seed=0):
    """4 patterns with n in {32, 90, 144, 300}, nnz ~ 14 n, and a random
    interleaving of their systems in one request list."""
    pats, As, bs = [], [], []
    for n in NS:
        pat, A, b = synth.batch("random", B_each, n=n, density=13.0 / (n - 1))
        pats.append(pat); As.append(A); bs.append(b)
    req_p = np.repeat(np.arange(len(NS)), B_each).astype(np.int32)
    req_i = np.tile(np.arange(B_each), len(NS)).astype(np.int32)
    perm = np.random.default_rng(seed).permutation(req_p.size)
    return pats, As, bs, req_p[perm], req_i[perm]


OPTS = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)


def per_pattern(P, A, b, **kw):
    x = torch.empty_like(b)
    out = bilu.solve_ex(P, A, b, x, bilu.opts(**kw))
    torch.cuda.synchronize()
    return x.cpu().numpy(), [t.cpu().numpy() for t in out[:5]]


@pytest.mark.parametrize("precond", [bilu.ILU0, bilu.ILU0_ISAI])
def test_uneven_matches_per_pattern_and_tuning(precond):
    print(f"case uneven precond={precond}", flush=True)
    B_each = 384
    pats, As, bs, req_p, req_i = c3(B_each)
    kw = dict(OPTS, precond=precond)
    P = [bilu.Pattern.from_synth(p) for p in pats]
    dA = [torch.from_numpy(A).to(dev()) for A in As]
    db = [torch.from_numpy(b).to(dev()) for b in bs]
    dx = [torch.empty_like(b) for b in db]
    plan = bilu.Uneven(P, req_p, req_i, bilu.opts(**kw))
    out = plan.outputs()
    plan.solve(dA, db, dx, out)
    torch.cuda.synchronize()
    first = [t.cpu().numpy().copy() for t in out]
    xs = [x.cpu().numpy().copy() for x in dx]
    for p in range(len(P)):
        xr, res = per_pattern(P[p], dA[p], db[p], **kw)
        assert np.array_equal(xs[p], xr), f"pattern {p}: x differs"
        rows = np.nonzero(req_p == p)[0]
        for name, u, v in zip(["iters", "conv", "iter_res", "true_res", "refine"], first, res):
            assert np.array_equal(u[rows], v[req_i[rows]]), f"pattern {p}: {name} differs"
    assert (first[1] == 1).all()
    # the autotuner picks the merge point and the workgroup sizes (alone, then
    # inside the whole launch); the results stay bit for bit
    plan.tune(dA, db, dx, out)
    for p in range(len(P)):
        fw, sw, jobs = plan.config(p)
        assert (fw == 0) == plan.merged(p) and sw in (4, 8, 12) and jobs == B_each
        assert fw in (0, 2, 4, 8)
        tab = plan.tune_table(p)
        # the record placement is a candidate for the ISAI templates of 48 / 72 slots (n = 300 here)
        place = precond == bilu.ILU0_ISAI and NS[p] == 300
        assert len(tab) == (24 if place else 12) and sum(c["merged"] for c in tab) == (6 if place else 3)
        assert sum(c["records_global"] for c in tab) == (12 if place else 0)
        assert all(c["alone_ms"] > 0 and c["corun_ms"] > 0 for c in tab)
        print(f"pattern n={NS[p]}: merged={plan.merged(p)} factor_waves={fw} solve_waves={sw} "
              f"records_global={plan.records_global(p)}", flush=True)
    for x in dx:
        x.zero_()
    plan.solve(dA, db, dx, out)
    torch.cuda.synchronize()
    for p in range(len(P)):
        assert np.array_equal(dx[p].cpu().numpy(), xs[p])
    for u, v in zip(first, out):
        assert np.array_equal(u, v.cpu().numpy())
    plan.close()


@pytest.mark.parametrize("precond", [bilu.ILU0, bilu.ILU0_ISAI])
def test_uneven_merged_plan(precond):
    """a plan fixed on the merged kernel (opts.merge): bitwise equal to the
    per-pattern segregated solves; the tuner keeps the merge point fixed"""
    print(f"case uneven merged precond={precond}", flush=True)
    pats, As, bs, req_p, req_i = c3(256, seed=9)
    P = [bilu.Pattern.from_synth(p) for p in pats]
    dA = [torch.from_numpy(A).to(dev()) for A in As]
    db = [torch.from_numpy(b).to(dev()) for b in bs]
    dx = [torch.empty_like(b) for b in db]
    plan = bilu.Uneven(P, req_p, req_i, bilu.opts(precond=precond, merge=bilu.MERGED, **OPTS))
    out = plan.outputs()
    for tuned in (False, True):
        if tuned:
            plan.tune(dA, db, dx, out)
        plan.solve(dA, db, dx, out)
        torch.cuda.synchronize()
        for p in range(len(P)):
            assert plan.merged(p)
            xr, res = per_pattern(P[p], dA[p], db[p], precond=precond, merge=bilu.SEGREGATED, **OPTS)
            assert np.array_equal(dx[p].cpu().numpy(), xr), f"pattern {p}: x differs"
            rows = np.nonzero(req_p == p)[0]
            for name, u, v in zip(["iters", "conv", "iter_res", "true_res", "refine"], out, res):
                assert np.array_equal(u.cpu().numpy()[rows], v[req_i[rows]]), f"pattern {p}: {name} differs"
    assert len(plan.tune_table(0)) == 3
    plan.close()


def test_uneven_records_global_n300_bitwise():
    """the ISAI solve kernel with its slot records in global memory (the
    placement the tuner may pick for 72-slot patterns): bitwise the oracle's
    and the LDS placement's results, merged and segregated"""
    print("case uneven n300 records in global memory", flush=True)
    pat, A, b = synth.batch("random", 96, n=300, density=13.0 / 299)
    P = bilu.Pattern.from_synth(pat)
    dA, db = torch.from_numpy(A).to(dev()), torch.from_numpy(b).to(dev())
    o = O.batch_solve_split(pat, A, b, precond=O.ILU0_ISAI, precision=O.PREC_FP32_IR,
                            o=O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, trsv_order=O.TRSV_AVAIL), threads=8)
    plan = bilu.Uneven([P], np.zeros(96, np.int32), np.arange(96, dtype=np.int32),
                       bilu.opts(precond=bilu.ILU0_ISAI, merge=bilu.MERGE_AUTO, **OPTS))
    out = plan.outputs()
    for merged, fw, sw, rg in ((False, 4, 8, True), (True, 0, 8, True), (False, 8, 4, False), (True, 0, 4, False)):
        dx = torch.zeros_like(db)
        for t in out:
            t.zero_()
        plan.configure(0, merged, fw, sw, rg)
        assert plan.records_global(0) == rg and plan.merged(0) == merged
        plan.solve([dA], [db], [dx], out)
        torch.cuda.synchronize()
        assert np.array_equal(dx.cpu().numpy(), o[0]), f"merged={merged} records_global={rg}: x differs"
        for u, v in zip(out, (o[1], o[2], o[3], o[4], o[5])):
            assert np.array_equal(u.cpu().numpy(), v), f"merged={merged} records_global={rg}"
    plan.close()


def test_uneven_graph_replays_and_recapture():
    print("case uneven graph replays", flush=True)
    pats, As, bs, req_p, req_i = c3(256, seed=3)
    P = [bilu.Pattern.from_synth(p) for p in pats]
    dA = [torch.from_numpy(A).to(dev()) for A in As]
    db = [torch.from_numpy(b).to(dev()) for b in bs]
    dx = [torch.empty_like(b) for b in db]
    plan = bilu.Uneven(P, req_p, req_i, bilu.opts(precond=bilu.ILU0, **OPTS))
    out = plan.outputs()
    ref = None
    for rep in range(4):
        for x in dx:
            x.fill_(float("nan"))          # eager work between the replays
        for t in out:
            t.zero_()
        torch.cuda.synchronize()
        plan.solve(dA, db, dx, out)
        torch.cuda.synchronize()
        got = [x.cpu().numpy().copy() for x in dx] + [t.cpu().numpy().copy() for t in out]
        if ref is None:
            ref = got
        else:
            for u, v in zip(ref, got):
                assert np.array_equal(u, v), f"replay {rep} differs"
    # new output buffers: the plan captures a new graph
    out2 = plan.outputs()
    dx2 = [torch.empty_like(b) for b in db]
    plan.solve(dA, db, dx2, out2)
    torch.cuda.synchronize()
    for u, v in zip(ref[:4], dx2):
        assert np.array_equal(u, v.cpu().numpy())
    assert np.array_equal(ref[4 + 1], out2[1].cpu().numpy())
    plan.close()


@pytest.mark.parametrize("precond", [bilu.ILU0, bilu.ILU0_ISAI])
def test_uneven_n300_against_oracle(precond):
    print(f"case uneven n300 oracle precond={precond}", flush=True)
    pat, A, b = synth.batch("random", 48, n=300, density=13.0 / 299)
    P = bilu.Pattern.from_synth(pat)
    dA, db = torch.from_numpy(A).to(dev()), torch.from_numpy(b).to(dev())
    dx = torch.empty_like(db)
    plan = bilu.Uneven([P], np.zeros(48, np.int32), np.arange(48, dtype=np.int32),
                       bilu.opts(precond=precond, **OPTS))
    out = plan.outputs()
    plan.solve([dA], [db], [dx], out)
    torch.cuda.synchronize()
    o = O.batch_solve_split(pat, A, b, precond=O.ILU0 if precond == bilu.ILU0 else O.ILU0_ISAI,
                            precision=O.PREC_FP32_IR,
                            o=O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, trsv_order=O.TRSV_AVAIL), threads=8)
    assert np.array_equal(dx.cpu().numpy(), o[0])
    for u, v in zip(out, (o[1], o[2], o[3], o[4], o[5])):
        assert np.array_equal(u.cpu().numpy(), v)
    assert (o[2] == 1).all()
    plan.close()


def test_uneven_rejects_bad_plans():
    pats, As, bs, req_p, req_i = c3(4)
    P = [bilu.Pattern.from_synth(p) for p in pats]
    with pytest.raises(bilu.BiluError):     # the same handle twice
        bilu.Uneven([P[0], P[0]], np.zeros(2, np.int32), np.arange(2, dtype=np.int32), bilu.opts(**OPTS))
    with pytest.raises(bilu.BiluError):     # FP64 is served by bilu_solve
        bilu.Uneven(P, req_p, req_i, bilu.opts(tol=1e-10, tol_mode=1))
    with pytest.raises(bilu.BiluError):     # request of an unknown pattern
        bilu.Uneven(P, np.full(2, 7, np.int32), np.arange(2, dtype=np.int32), bilu.opts(**OPTS))


def test_uneven_parilu_stop_and_empty_pattern():
    """ParILU with the residual stop through the plan (bitwise equal to the
    per-pattern solve), with one pattern that has no requests."""
    print("case uneven parilu", flush=True)
    pats, As, bs, req_p, req_i = c3(128, seed=5)
    keep = req_p != 1                      # pattern n = 90 gets no requests
    req_p, req_i = req_p[keep], req_i[keep]
    kw = dict(OPTS, precond=bilu.PARILU0, parilu_sweeps=40, parilu_stop_tol=1e-8)
    P = [bilu.Pattern.from_synth(p) for p in pats]
    dA = [torch.from_numpy(A).to(dev()) for A in As]
    db = [torch.from_numpy(b).to(dev()) for b in bs]
    dx = [torch.zeros_like(b) for b in db]
    plan = bilu.Uneven(P, req_p, req_i, bilu.opts(**kw))
    assert plan.config(1)[2] == 0
    out = plan.outputs()
    plan.solve(dA, db, dx, out)
    torch.cuda.synchronize()
    for p in (0, 2, 3):
        xr, res = per_pattern(P[p], dA[p], db[p], **kw)
        assert np.array_equal(dx[p].cpu().numpy(), xr)
        rows = np.nonzero(req_p == p)[0]
        assert np.array_equal(out[0].cpu().numpy()[rows], res[0][req_i[rows]])
        assert np.array_equal(out[1].cpu().numpy()[rows], res[1][req_i[rows]])
    assert (out[1].cpu().numpy() == 1).all()
    assert (dx[1].cpu().numpy() == 0).all()     # untouched
    plan.close()
    # an empty request list is a no-op
    plan0 = bilu.Uneven(P, np.zeros(0, np.int32), np.zeros(0, np.int32), bilu.opts(**kw))
    plan0.solve(dA, db, dx, plan0.outputs())
    plan0.close()
