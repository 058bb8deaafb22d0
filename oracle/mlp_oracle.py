"""numpy restatement of the learned warm start's MLP (test infrastructure:
the checker of the library's BF16 MFMA forward / backward kernels).

Same model and scaling as the library: x = asinh(a / 1e-8) / 24, two
tanh-GELU hidden layers, y = h2 W3^T + b3, loss = mean((y - t)^2) over the
pattern entries with t = asinh(F / 1e-8) / 24.  `bf16=True` rounds the
operands the kernels hold in BF16 (features, weights, hidden activations,
backpropagated gradients) with round-to-nearest-even; accumulation is in
float64 here (FP32 in the kernels).  The gradient formulas are pinned by
central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

SIGMA, YS = 1e-8, 24.0
C0, C1 = 0.7978845608028654, 0.044715


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + C1 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C0 * (z + C1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C0 * (1.0 + 3.0 * C1 * z * z)


def unpack(w, nnz, hidden):
    """Packed layout: W1[h][nnz], b1[h], W2[h][h], b2[h], W3[nnz][h], b3[nnz]."""
    sizes = [hidden * nnz, hidden, hidden * hidden, hidden, nnz * hidden, nnz]
    parts, o = [], 0
    for s in sizes:
        parts.append(np.asarray(w[o:o + s], dtype=np.float64))
        o += s
    W1, b1, W2, b2, W3, b3 = parts
    return W1.reshape(hidden, nnz), b1, W2.reshape(hidden, hidden), b2, W3.reshape(nnz, hidden), b3


def features(A):
    return np.arcsinh(np.asarray(A, dtype=np.float64) / SIGMA) / YS


def targets(F):
    return np.arcsinh(np.asarray(F, dtype=np.float64) / SIGMA) / YS


def base(pat, A):
    """Base guess in the scaled space: s = a_ij on / above the diagonal,
    a_ij / a_jj below it."""
    A = np.asarray(A, dtype=np.float64)
    s = A.copy()
    low = pat.col_idx < pat.rows
    s[..., low] = A[..., low] / A[..., pat.diag[pat.col_idx[low]]]
    return np.arcsinh(s / SIGMA) / YS


def forward(w, A, nnz, hidden, bf=False):
    r = bf16 if bf else (lambda v: v)
    W1, b1, W2, b2, W3, b3 = unpack(w, nnz, hidden)
    X = r(features(A))
    Z1 = X @ r(W1).T + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ r(W2).T + b2
    H2 = r(gelu(Z2))
    Y = H2 @ r(W3).T + b3
    return Y, (X, Z1, H1, Z2, H2)


def predict(w, pat, A, hidden, bf=False):
    Y, _ = forward(w, A, pat.nnz, hidden, bf)
    return SIGMA * np.sinh(YS * (base(pat, A) + Y))


def loss_and_grad(w, pat, A, F, hidden, bf=False):
    """Mean squared error of the network output against the scaled factor
    minus the base guess, and its gradient (packed layout)."""
    nnz = pat.nnz
    r = bf16 if bf else (lambda v: v)
    W1, b1, W2, b2, W3, b3 = unpack(w, nnz, hidden)
    Y, (X, Z1, H1, Z2, H2) = forward(w, A, nnz, hidden, bf)
    T = targets(F) - base(pat, A)
    B = A.shape[0]
    loss = np.mean((Y - T) ** 2)
    dY = r(2.0 * (Y - T) / (B * nnz))
    gW3 = dY.T @ H2
    gb3 = dY.sum(0)
    dZ2 = r((dY @ r(W3)) * gelu_grad(Z2))
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = r((dZ2 @ r(W2)) * gelu_grad(Z1))
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
