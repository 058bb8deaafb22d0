"""GPU checks of the learned warm start: the BF16 MFMA forward / backward
kernels against the numpy MLP oracle (BF16 operand rounding emulated,
accumulation order differs: tolerances below), the weight exchange, and
what the trained network gains for ParILU(0)."""
import numpy as np
import pytest

import bilu
import mlp_oracle as MO
import oracle as O
import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def exact_factors(P, dA):
    F = torch.empty_like(dA)
    bilu.ilu0(P, dA, F)
    return F


def test_mlp_weights_round_trip():
    print("case mlp weights", flush=True)
    pat, A, b = synth.batch("random", 4, n=144)
    P = bilu.Pattern.from_synth(pat)
    net = bilu.Mlp(P, hidden=256, seed=3)
    w = net.get_weights()
    assert w.numel() == 256 * pat.nnz * 2 + 256 * 256 + 2 * 256 + pat.nnz
    w2 = torch.randn_like(w) * 0.01
    net.set_weights(w2)
    assert torch.equal(net.get_weights(), w2)


def test_mlp_gradient_matches_oracle():
    print("case mlp gradient", flush=True)
    pat, A, b = synth.batch("random", 200, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA = to_dev(A)
    dF = exact_factors(P, dA)
    net = bilu.Mlp(P, hidden=512, seed=5)
    w0 = net.get_weights()
    net.set_weights(w0 + 0.02 * torch.randn_like(w0))   # nonzero output layer: every block gets a gradient
    loss, g = net.gradient(dA, dF)
    w = net.get_weights().cpu().numpy()
    torch.cuda.synchronize()
    lo, go = MO.loss_and_grad(w, pat, A, dF.cpu().numpy(), 512, bf=True)
    gg = g.cpu().numpy().astype(np.float64)
    assert loss.item() == pytest.approx(lo, rel=2e-3)
    rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
    cos = gg @ go / (np.linalg.norm(gg) * np.linalg.norm(go))
    print(f"gradient: relative error {rel:.3e}, cosine {cos:.6f}", flush=True)
    assert rel < 2e-2 and cos > 0.9998
    # per layer as well (a transposed operand would spoil one block)
    o = 0
    for size in (512 * pat.nnz, 512, 512 * 512, 512, pat.nnz * 512, pat.nnz):
        blk = slice(o, o + size)
        assert np.linalg.norm(gg[blk] - go[blk]) <= 3e-2 * np.linalg.norm(go[blk]) + 1e-12
        o += size


def test_mlp_forward_matches_oracle():
    print("case mlp forward", flush=True)
    pat, A, b = synth.batch("random", 300, n=144)
    P = bilu.Pattern.from_synth(pat)
    net = bilu.Mlp(P, hidden=512, seed=9)
    w0 = net.get_weights()
    net.set_weights(w0 + 0.02 * torch.randn_like(w0))
    w = net.get_weights().cpu().numpy()
    Fg = net.infer(to_dev(A)).cpu().numpy()
    Y = MO.forward(w, A, pat.nnz, 512, bf=True)[0] + MO.base(pat, A)
    Yg = MO.targets(Fg)
    assert np.abs(Yg - Y).max() < 1e-2 * max(1.0, np.abs(Y).max())


def test_mlp_training_gain():
    """Train on exact factors of fresh systems (16384 here; the benchmark
    trains on 262144 for 30 epochs), then count in-place ParILU(0) sweeps to
    a factor residual of 1e-8 on held-out systems from three starts: A (no
    warm start), the untrained network (its base guess a_ij, a_ij / a_jj)
    and the trained network.  What is asserted: the warm start saves sweeps
    against A, training lowers the loss, the trained guess stays within a
    small margin of its base guess, and every solve stays exact.  What is
    reported: trained against base -- on these i.i.d. systems the network
    does not beat its own base guess (the base guess is already within
    1e-4 relative on most entries; even exact values on every entry that
    the base guess gets wrong by 0.1 % save only ~0.1 sweep)."""
    print("case mlp training gain", flush=True)
    P_pat, A, b = synth.batch("random", 16384, first=1 << 20, n=144)
    P = bilu.Pattern.from_synth(P_pat)
    dA = to_dev(A)
    dF = exact_factors(P, dA)
    net = bilu.Mlp(P, hidden=512, seed=11)
    net0 = bilu.Mlp(P, hidden=512, seed=11)
    loss = torch.zeros(1, dtype=torch.float32, device=dev())
    losses = []
    mb = 1024
    for step in range(600):
        s0 = (step * mb) % A.shape[0]
        net.train_step(dA[s0:s0 + mb], dF[s0:s0 + mb], 3e-4 if step < 400 else 1e-4, loss)
        if step % 100 == 0:
            losses.append(loss.item())
    torch.cuda.synchronize()
    print(f"training loss {losses}", flush=True)
    assert losses[-1] < losses[0]
    pat, At, bt = synth.batch("random", 2048, n=144)        # held out
    dAt, dbt = to_dev(At), to_dev(bt)
    kw = dict(precond=bilu.PARILU0_ISAI, parilu_sweeps=60, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR, parilu_order=bilu.PARILU_INPLACE)
    o = bilu.opts(**kw)
    x = torch.empty_like(dbt)
    bn = np.linalg.norm(bt, axis=1)

    def check(r):
        torch.cuda.synchronize()
        assert (r[1].cpu().numpy() == 1).all() and (r[3].cpu().numpy() / bn).max() <= 1e-10
        fs = r[5].cpu().numpy()
        assert (fs > 0).all()
        return fs.mean()
    s_a = check(bilu.solve_ex(P, dAt, dbt, x, o, factor_status=torch.empty(2048, dtype=torch.int32, device=dev())))
    s_base = check(net0.solve(P, dAt, dbt, x, o))
    s_net = check(net.solve(P, dAt, dbt, x, o))
    print(f"in-place ParILU sweeps to 1e-8: from A {s_a:.3f}, base guess {s_base:.3f}, trained network "
          f"{s_net:.3f}", flush=True)
    assert s_base < s_a
    assert s_net < s_base + 0.25
    # the scaled error of the guess: trained network vs base guess vs A (reported)
    guess = net.infer(dAt)
    Fe = exact_factors(P, dAt).cpu().numpy()
    eg = np.abs(MO.targets(guess.cpu().numpy()) - MO.targets(Fe)).mean()
    eb = np.abs(MO.base(pat, At) - MO.targets(Fe)).mean()
    ea = np.abs(MO.targets(At) - MO.targets(Fe)).mean()
    print(f"mean scaled error: trained guess {eg:.4f}, base guess {eb:.4f}, A {ea:.4f}", flush=True)
    assert eg < ea
    net.close()
    net0.close()


def _fused_vs_split(P, net, dA, db, precond, sweeps=40, stop_tol=1e-8, order=bilu.PARILU_JACOBI):
    """bilu_mlp_solve against bilu_mlp_infer + bilu_solve_ex(F_init): every
    output bit-equal (the fused warm start runs the same FP32 operations)."""
    o = bilu.opts(precond=precond, parilu_sweeps=sweeps, parilu_stop_tol=stop_tol, tol=1e-10,
                  tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR, parilu_order=order)
    x1, x2 = torch.empty_like(db), torch.empty_like(db)
    r1 = net.solve(P, dA, db, x1, o)
    r2 = bilu.solve_ex(P, dA, db, x2, o, F_init=net.infer(dA))
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)
    for name, a, b in zip(("iters", "conv", "iter_res", "true_res", "refine", "factor_status"), r1, r2):
        assert torch.equal(a, b), name
    return r1


@pytest.mark.parametrize("precond", ["PARILU0", "PARILU0_ISAI"])
@pytest.mark.parametrize("order", ["PARILU_JACOBI", "PARILU_INPLACE"])
def test_mlp_solve_fused_matches_split(precond, order):
    print(f"case mlp fused solve {precond} {order}", flush=True)
    pat, A, b = synth.batch("random", 768, first=3 << 20, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    net = bilu.Mlp(P, hidden=256, seed=5)
    # a few steps so the guess is not the untrained one
    dF = exact_factors(P, dA)
    for step in range(20):
        net.train_step(dA[:512], dF[:512], 3e-4)
    it, cv, _, tr, _, fs = _fused_vs_split(P, net, dA, db, getattr(bilu, precond), order=getattr(bilu, order))
    assert (cv.cpu().numpy() == 1).all()
    assert (tr.cpu().numpy() / np.linalg.norm(b, axis=1)).max() <= 1e-10


def test_mlp_solve_fused_chunks():
    """more systems than one network pass (65536): the fused solve walks the
    batch in chunks; a small pattern keeps the case cheap"""
    print("case mlp fused solve chunks", flush=True)
    pat = synth.stencil5_pattern(4, 4)
    B = 65536 + 37
    A = synth.stencil5_values(pat, np.arange(B))
    b = synth.rhs(pat.n, np.arange(B))
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    net = bilu.Mlp(P, hidden=128, seed=3)
    _fused_vs_split(P, net, dA, db, bilu.PARILU0, sweeps=12)


def test_mlp_solve_rejects():
    pat, A, b = synth.batch("random", 4, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    net = bilu.Mlp(P, hidden=128, seed=1)
    x = torch.empty_like(db)
    for kw in (dict(precond=bilu.ILU0, precision=bilu.PRECISION_FP32_IR),
               dict(precond=bilu.PARILU0, precision=bilu.PRECISION_FP64)):
        with pytest.raises(bilu.BiluError):
            net.solve(P, dA, db, x, bilu.opts(**kw))
    pat2 = synth.stencil5_pattern(4, 4)
    P2 = bilu.Pattern.from_synth(pat2)
    A2 = synth.stencil5_values(pat2, np.arange(4))
    with pytest.raises(bilu.BiluError):
        net.solve(P2, to_dev(A2), to_dev(synth.rhs(pat2.n, np.arange(4))), torch.empty(4, pat2.n, dtype=torch.float64,
                  device=dev()), bilu.opts(precond=bilu.PARILU0, precision=bilu.PRECISION_FP32_IR))


@pytest.mark.parametrize("precond", ["PARILU0", "PARILU0_ISAI"])
@pytest.mark.parametrize("order", ["PARILU_JACOBI", "PARILU_INPLACE"])
def test_mlp_solve_merged_matches_segregated(precond, order):
    """bilu_mlp_solve with the merged kernel (warm start, sweeps, ISAI and the
    solve in one launch) equals the segregated launches bit for bit"""
    print(f"case mlp merged solve {precond} {order}", flush=True)
    pat, A, b = synth.batch("random", 640, first=5 << 20, n=144)
    P = bilu.Pattern.from_synth(pat)
    dA, db = to_dev(A), to_dev(b)
    net = bilu.Mlp(P, hidden=256, seed=9)
    dF = exact_factors(P, dA)
    for step in range(10):
        net.train_step(dA[:512], dF[:512], 3e-4)
    res = {}
    for merge in (bilu.MERGED, bilu.SEGREGATED):
        o = bilu.opts(precond=getattr(bilu, precond), parilu_sweeps=40, parilu_stop_tol=1e-8, tol=1e-10,
                      tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR,
                      parilu_order=getattr(bilu, order), merge=merge)
        x = torch.empty_like(db)
        r = net.solve(P, dA, db, x, o)
        torch.cuda.synchronize()
        res[merge] = (x,) + tuple(r)
    for name, u, v in zip(("x", "iters", "conv", "iter_res", "true_res", "refine", "factor_status"),
                          res[bilu.MERGED], res[bilu.SEGREGATED]):
        assert torch.equal(u, v), name
    assert (res[bilu.MERGED][2].cpu().numpy() == 1).all()
    net.close()
