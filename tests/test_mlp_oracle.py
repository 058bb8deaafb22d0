"""The numpy MLP restatement (checker of the BF16 MFMA kernels): its
backpropagation is pinned by central finite differences, its GELU
derivative likewise, its bfloat16 rounding by exact cases."""
import numpy as np
import pytest

import mlp_oracle as M
import synth


def test_bf16_rounding_exact_cases():
    one = 1.0
    assert M.bf16(one) == 1.0
    # 1 + 2^-8 is the midpoint between 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 is the midpoint between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.all(np.abs(M.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)


def test_backprop_matches_finite_differences():
    pat, A, _ = synth.batch("random", 3, n=6, density=0.5)
    import oracle as O
    F = np.stack([O.ilu0(pat, A[s])[0] for s in range(3)])
    nnz, hid = pat.nnz, 5
    rng = np.random.default_rng(2)
    w = rng.standard_normal(hid * nnz + hid + hid * hid + hid + nnz * hid + nnz) * 0.3
    loss, g = M.loss_and_grad(w, pat, A, F, hid)
    h = 1e-6
    fd = np.empty_like(w)
    for i in range(w.size):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd[i] = (M.loss_and_grad(wp, pat, A, F, hid)[0] - M.loss_and_grad(wm, pat, A, F, hid)[0]) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-9)
    assert loss == pytest.approx(np.mean((M.forward(w, A, nnz, hid)[0] - (M.targets(F) - M.base(pat, A))) ** 2))


def test_base_guess_is_the_first_jacobi_guess():
    # below the diagonal a_ij / a_jj, on and above a_ij: for a closed
    # 2x2 pattern these are the exact factors' first row and multiplier
    pat = synth.Pattern(2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]))
    A = np.array([4.0, 1.0, 2.0, 3.0])
    s = M.SIGMA * np.sinh(M.YS * M.base(pat, A))
    assert np.allclose(s, [4.0, 1.0, 0.5, 3.0], rtol=1e-12)


def test_scaling_round_trip():
    a = np.array([-1e2, -3.0, -1e-6, 0.0, 2e-7, 1.0, 5e2])
    x = M.features(a)
    assert np.allclose(M.SIGMA * np.sinh(M.YS * x), a, rtol=1e-12, atol=1e-20)
    assert np.all(np.sign(x) == np.sign(a))
