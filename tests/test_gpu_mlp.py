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
    """Train on exact factors of fresh systems (16384 here; the benchmark's
    learned mode trains on 262144), then ParILU(0) sweeps to a factor
    residual of 1e-8 from the network's guess against the standard start
    (A): the warm start must save sweeps, and the solves stay exact."""
    print("case mlp training gain", flush=True)
    P_pat, A, b = synth.batch("random", 16384, first=1 << 20, n=144)
    P = bilu.Pattern.from_synth(P_pat)
    dA = to_dev(A)
    dF = exact_factors(P, dA)
    net = bilu.Mlp(P, hidden=512, seed=11)
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
    pat, At, bt = synth.batch("random", 512, n=144)        # held out
    dAt, dbt = to_dev(At), to_dev(bt)
    guess = net.infer(dAt)
    kw = dict(precond=bilu.PARILU0, parilu_sweeps=60, parilu_stop_tol=1e-8, tol=1e-10,
              tol_mode=bilu.TOL_RELATIVE, precision=bilu.PRECISION_FP32_IR)
    x = torch.empty_like(dbt)
    _, cv1, _, tr1, _, fs1 = bilu.solve_ex(P, dAt, dbt, x, bilu.opts(**kw), F_init=guess)
    _, cv0, _, tr0, _, fs0 = bilu.solve_ex(P, dAt, dbt, x, bilu.opts(**kw))
    torch.cuda.synchronize()
    s1, s0 = fs1.cpu().numpy(), fs0.cpu().numpy()
    print(f"ParILU sweeps to 1e-8: from A {s0.mean():.2f}, from the network {s1.mean():.2f}", flush=True)
    assert (s0 > 0).all() and (s1 >= 0).all()
    assert s1.mean() < s0.mean()
    bn = np.linalg.norm(bt, axis=1)
    assert (cv1.cpu().numpy() == 1).all() and (tr1.cpu().numpy() / bn).max() <= 1e-10
    # the guess is closer to the exact factors than the standard start (A);
    # its error against the untrained network's base guess is reported
    Fe = exact_factors(P, dAt).cpu().numpy()
    eg = np.abs(MO.targets(guess.cpu().numpy()) - MO.targets(Fe)).mean()
    eb = np.abs(MO.base(pat, At) - MO.targets(Fe)).mean()
    ea = np.abs(MO.targets(At) - MO.targets(Fe)).mean()
    print(f"mean scaled error: trained guess {eg:.4f}, base guess {eb:.4f}, A {ea:.4f}", flush=True)
    assert eg < ea
