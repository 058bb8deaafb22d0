"""Synthetic batches of small sparse systems sharing one CSR pattern.

Counter-based generation: every random number is a pure function of
(seed, stream, system id, entry id) through splitmix64, so a batch (or any
prefix or slice of it) is bit-identical however and wherever it is generated.

Workloads (sizes are the ones the solver is benchmarked and tested on):

* ``stencil5``: 10x10 five-point grid (n=100, nnz=460); off-diagonals
  -U(0.5,1.5), diagonal sum|offdiag| * U(1.1,2.0), rhs U(-1,1).
* ``random_pattern``: n rows, full diagonal, symmetric off-diagonal pattern
  with pair density d (0.09 at n=144 gives nnz ~ 2000; 13/(n-1) gives
  nnz ~ 14 n); off-diagonal magnitudes log-uniform 1e-6..1e2 with random
  sign, diagonal sum|offdiag| * U(1.05,2.0), rhs U(-1,1).

The reference solver reads these systems from Matrix Market directories
(ReadWriteData.cpp:258-464); all pages share one pattern (:385-429) with
sorted columns (:239-253) and page-major values (:433).  The generated CSR
follows the same layout: int32 row_ptrs[n+1], int32 col_idx[nnz] sorted per
row, float64 values[batch*nnz], float64 rhs[batch*n].
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)

STREAM_PATTERN = 1
STREAM_MAG = 2
STREAM_SIGN = 3
STREAM_DIAG = 4
STREAM_RHS = 5

PATTERN_SEED = 144
VALUE_SEED = 20261203


def _splitmix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = z + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        return z ^ (z >> np.uint64(31))


def u01(seed: int, stream: int, system, entry) -> np.ndarray:
    """Uniform [0,1) doubles keyed on (seed, stream, system, entry)."""
    s = np.asarray(system, dtype=np.uint64)
    e = np.asarray(entry, dtype=np.uint64)
    k0 = _splitmix64(np.uint64(seed) ^ (np.uint64(stream) << np.uint64(56)))
    h = _splitmix64(_splitmix64(k0 ^ s) ^ e)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Pattern:
    """Shared CSR pattern (sorted columns, full diagonal)."""

    def __init__(self, n: int, row_ptrs: np.ndarray, col_idx: np.ndarray, name: str = ""):
        self.n = int(n)
        self.row_ptrs = np.ascontiguousarray(row_ptrs, dtype=np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.nnz = int(self.row_ptrs[-1])
        self.name = name
        rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(self.row_ptrs))
        self.rows = rows
        self.diag = np.full(n, -1, dtype=np.int32)
        on = rows == self.col_idx
        self.diag[rows[on]] = np.nonzero(on)[0].astype(np.int32)

    def dense(self, vals: np.ndarray) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        A[self.rows, self.col_idx] = vals
        return A


def stencil5_pattern(nx: int = 10, ny: int = 10) -> Pattern:
    n = nx * ny
    rp = [0]
    ci = []
    for i in range(n):
        gx, gy = i % nx, i // nx
        cols = []
        if gy > 0:
            cols.append(i - nx)
        if gx > 0:
            cols.append(i - 1)
        cols.append(i)
        if gx < nx - 1:
            cols.append(i + 1)
        if gy < ny - 1:
            cols.append(i + nx)
        ci.extend(cols)
        rp.append(len(ci))
    return Pattern(n, np.array(rp), np.array(ci), name=f"stencil5_{nx}x{ny}")


def random_pattern(n: int, density: float, seed: int = PATTERN_SEED) -> Pattern:
    """Full diagonal plus (i,j),(j,i) for i<j iff U01(seed, i*n+j) < density."""
    iu, ju = np.triu_indices(n, 1)
    keep = u01(seed, STREAM_PATTERN, 0, iu.astype(np.uint64) * np.uint64(n) + ju.astype(np.uint64)) < density
    M = np.eye(n, dtype=bool)
    M[iu[keep], ju[keep]] = True
    M[ju[keep], iu[keep]] = True
    rows, cols = np.nonzero(M)  # row-major, sorted columns
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    return Pattern(n, rp, cols, name=f"random_n{n}_d{density:.4g}_s{seed}")


def _values_common(pat: Pattern, systems: np.ndarray, seed: int, offdiag_fn, diag_lo, diag_hi):
    nnz, n = pat.nnz, pat.n
    B = systems.size
    e = np.arange(nnz, dtype=np.uint64)
    off = pat.rows != pat.col_idx
    vals = offdiag_fn(systems[:, None], e[None, :])
    vals[:, ~off] = 0.0
    rowsum = np.zeros((B, n))
    np.add.at(rowsum, (slice(None), pat.rows), np.abs(vals))
    dfac = diag_lo + (diag_hi - diag_lo) * u01(seed, STREAM_DIAG, systems[:, None], np.arange(n, dtype=np.uint64)[None, :])
    vals[:, pat.diag] = rowsum * dfac
    return vals


def stencil5_values(pat: Pattern, systems, seed: int = VALUE_SEED) -> np.ndarray:
    systems = np.asarray(systems, dtype=np.uint64).reshape(-1)

    def off(s, e):
        return -(0.5 + u01(seed, STREAM_MAG, s, e))

    return _values_common(pat, systems, seed, off, 1.1, 2.0)


def random_values(pat: Pattern, systems, seed: int = VALUE_SEED) -> np.ndarray:
    systems = np.asarray(systems, dtype=np.uint64).reshape(-1)

    def off(s, e):
        mag = np.power(10.0, -6.0 + 8.0 * u01(seed, STREAM_MAG, s, e))
        sgn = np.where(u01(seed, STREAM_SIGN, s, e) < 0.5, -1.0, 1.0)
        return mag * sgn

    return _values_common(pat, systems, seed, off, 1.05, 2.0)


def rhs(n: int, systems, seed: int = VALUE_SEED) -> np.ndarray:
    systems = np.asarray(systems, dtype=np.uint64).reshape(-1)
    return -1.0 + 2.0 * u01(seed, STREAM_RHS, systems[:, None], np.arange(n, dtype=np.uint64)[None, :])


def batch(kind: str, B: int, first: int = 0, n: int = 144, density: float | None = None,
          pattern_seed: int = PATTERN_SEED, value_seed: int = VALUE_SEED, chunk: int = 4096):
    """Return (pattern, values[B,nnz], rhs[B,n]) for systems first..first+B-1."""
    if kind == "stencil5":
        pat = stencil5_pattern()
        vf = stencil5_values
    elif kind == "random":
        if density is None:
            density = 0.09 if n == 144 else 13.0 / (n - 1)
        pat = random_pattern(n, density, pattern_seed)
        vf = random_values
    else:
        raise ValueError(kind)
    vals = np.empty((B, pat.nnz))
    b = np.empty((B, pat.n))
    for c0 in range(0, B, chunk):
        c1 = min(B, c0 + chunk)
        ids = np.arange(first + c0, first + c1, dtype=np.uint64)
        vals[c0:c1] = vf(pat, ids, value_seed)
        b[c0:c1] = rhs(pat.n, ids, value_seed)
    return pat, vals, b
