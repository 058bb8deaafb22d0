"""ctypes wrapper of the CPU oracle (liboracle.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it, as the checker or the timed CPU baseline.  The
solver product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

CONVERGED, NOT_CONVERGED, ZERO_PIVOT, NONFINITE = 1, -1, -2, -3
PC_NONE, PC_LU, PC_ISAI_LU, PC_ISAI_GEN = 0, 1, 2, 3
ISAI_GENERAL, ISAI_LOWER, ISAI_UPPER = 0, 1, 2
TOL_ABSOLUTE, TOL_RELATIVE = 0, 1
# batch precond codes of orc_batch_solve
ILU0, PARILU0, ILU0_ISAI, PARILU0_ISAI, GENERAL_ISAI, NONE = 0, 1, 2, 3, 4, 5


class Opts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("tol_mode", C.c_int),
        ("max_iter", C.c_int),
        ("dot_width", C.c_int),
        ("spmv_width", C.c_int),
        ("true_res_restart", C.c_int),
        ("max_restarts", C.c_int),
        ("trsv_order", C.c_int),
        ("inner_adapt", C.c_double),
    ]


TRSV_CSR, TRSV_AVAIL = 0, 1


def opts(tol=1e-11, tol_mode=TOL_ABSOLUTE, max_iter=500, dot_width=64, spmv_width=1,
         true_res_restart=0, max_restarts=3, trsv_order=TRSV_CSR, inner_adapt=0.5) -> Opts:
    """Defaults: reference tolerance semantics (ATOL=1e-11, MAX_ITER=500,
    PrecondBiCGSTAB.h:10-12) with the wave64 reduction order.  trsv_order 1
    is the row-sum order of the persistent kernel.  inner_adapt: the FP32
    inner solves of the mixed path stop at the relative residual
    max(inner_tol, inner_adapt * tol / |r|) (0: inner_tol)."""
    return Opts(tol, tol_mode, max_iter, dot_width, spmv_width, true_res_restart, max_restarts,
                trsv_order, inner_adapt)


def reference_opts(**kw) -> Opts:
    """Reference reduction order: warp-32 dots and warp-32 SpMV rows."""
    kw.setdefault("dot_width", 32)
    kw.setdefault("spmv_width", 32)
    return opts(**kw)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = np.ctypeslib.ndpointer
        i32 = P(np.int32, flags="C")
        f64 = P(np.float64, flags="C")
        f32 = P(np.float32, flags="C")
        L.orc_find_diag.argtypes = [C.c_int, i32, i32, i32]
        L.orc_isai_pattern.argtypes = [C.c_int, i32, i32, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_isai_on.argtypes = [C.c_int, i32, i32, f64, C.c_int, i32, i32, f64]
        L.orc_solve_isai_pow.argtypes = [C.c_int, i32, i32, f64, f64, f64, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(Opts), C.POINTER(C.c_float), C.POINTER(C.c_int),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_add_diagonal.argtypes = [C.c_int, i32, i32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_expand_values.argtypes = [C.c_int, C.c_int, C.c_int, i32, f64, f64]
        L.orc_dot.argtypes = [C.c_int, f64, f64, C.c_int]
        L.orc_dot.restype = C.c_double
        L.orc_spmv.argtypes = [C.c_int, i32, i32, f64, f64, f64, C.c_int]
        L.orc_ilu0.argtypes = [C.c_int, i32, i32, i32, f64]
        L.orc_parilu0_sweeps.argtypes = [C.c_int, i32, i32, i32, f64, f64, C.c_int]
        L.orc_factor_residual.argtypes = [C.c_int, i32, i32, i32, f64, f64]
        L.orc_factor_residual.restype = C.c_double
        L.orc_lu_solve.argtypes = [C.c_int, i32, i32, i32, f64, f64, f64, f64]
        L.orc_isai.argtypes = [C.c_int, i32, i32, i32, f64, C.c_int, f64]
        L.orc_trsv_levels.argtypes = [C.c_int, i32, i32, i32, i32, i32]
        L.orc_lu_solve_avail.argtypes = [C.c_int, i32, i32, i32, i32, i32, f64, f64, f64, f64]
        L.orc_bicgstab.argtypes = [C.c_int, i32, i32, i32, f64, f64, f64, C.c_int,
                                   C.c_void_p, C.c_void_p, C.POINTER(Opts),
                                   C.POINTER(C.c_float), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_batch_solve.argtypes = [C.c_int, C.c_int, i32, i32, C.c_int, f64, f64, f64,
                                      C.c_int, C.c_int, C.POINTER(Opts), f32,
                                      P(np.int32, flags="C"), f64, f64, C.c_int]
        L.orc_entry_order.argtypes = [C.c_int, i32, i32, i32, i32]
        L.orc_sweep_order.argtypes = [C.c_int, i32, i32, i32, i32]
        L.orc_isai_lu_f32.argtypes = [C.c_int, i32, i32, i32, f64, f32]
        L.orc_parilu0_stop.argtypes = [C.c_int, i32, i32, i32, f64, C.c_void_p, f64, C.c_int,
                                       C.c_double, C.POINTER(C.c_int), C.c_int]
        L.orc_batch_solve_split.argtypes = [C.c_int, C.c_int, i32, i32, C.c_int, f64, f64, f64,
                                            C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int,
                                            C.c_double, C.c_int, C.POINTER(Opts), f32, i32, f64,
                                            f64, i32, i32, C.c_int, C.c_int]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def find_diag(n, rp, ci):
    d = np.empty(n, dtype=np.int32)
    missing = lib().orc_find_diag(n, _i32(rp), _i32(ci), d)
    return d, missing


def isai_pattern(n, rp, ci, typ, power):
    """(row_ptrs, col_idx) of |P|^power, P the typed view's pattern
    (LOWER: columns < i plus i; UPPER: columns >= i; GENERAL: all)"""
    rp, ci = _i32(rp), _i32(ci)
    nnz = lib().orc_isai_pattern(n, rp, ci, typ, power, None, None)
    mrp, mci = np.empty(n + 1, np.int32), np.empty(nnz, np.int32)
    lib().orc_isai_pattern(n, rp, ci, typ, power, mrp.ctypes.data, mci.ctypes.data)
    return mrp, mci


def isai_on(pat, vals, typ, mrp, mci):
    """ISAI of the typed view of (pat, vals) on the pattern (mrp, mci): (values, status)"""
    out = np.zeros(int(mrp[-1]))
    st = lib().orc_isai_on(pat.n, _i32(pat.row_ptrs), _i32(pat.col_idx), _f64(vals), typ, _i32(mrp),
                           _i32(mci), out)
    return out, st


def solve_isai_pow(pat, A, b, precond=ILU0_ISAI, sweeps=5, power=2, o: Opts | None = None):
    """batch loop over orc_solve_isai_pow: (x, iter, conv, iter_res, true_res)"""
    o = o or opts()
    A, b = np.atleast_2d(_f64(A)), np.atleast_2d(_f64(b))
    B = A.shape[0]
    x = np.zeros((B, pat.n))
    it, cv = np.zeros(B, np.float32), np.zeros(B, np.int32)
    ir, tr = np.zeros(B), np.zeros(B)
    rp, ci = _i32(pat.row_ptrs), _i32(pat.col_idx)
    for s in range(B):
        fi, ii, di, dt = C.c_float(), C.c_int(), C.c_double(), C.c_double()
        xs = np.zeros(pat.n)
        rc = lib().orc_solve_isai_pow(pat.n, rp, ci, _f64(A[s]), _f64(b[s]), xs, precond, sweeps, power,
                                      C.byref(o), C.byref(fi), C.byref(ii), C.byref(di), C.byref(dt))
        if rc != 0:
            raise ValueError(f"orc_solve_isai_pow: {rc}")
        x[s], it[s], cv[s], ir[s], tr[s] = xs, fi.value, ii.value, di.value, dt.value
    return x, it, cv, ir, tr


def add_diagonal(n, rp, ci):
    """(row_ptrs, col_idx, holder) of the pattern with every missing (i,i)
    inserted; holder[k] = old position or -(i+1) for an inserted diagonal"""
    rp, ci = _i32(rp), _i32(ci)
    missing = lib().orc_add_diagonal(n, rp, ci, None, None, None)
    rpo = np.empty(n + 1, np.int32)
    cio = np.empty(len(ci) + missing, np.int32)
    h = np.empty(len(ci) + missing, np.int32)
    lib().orc_add_diagonal(n, rp, ci, rpo.ctypes.data, cio.ctypes.data, h.ctypes.data)
    return rpo, cio, h


def expand_values(holder, vals):
    """[batch, nnz] values on the augmented pattern (0 at inserted diagonals)"""
    vals = _f64(np.atleast_2d(vals))
    out = np.empty((vals.shape[0], len(holder)))
    lib().orc_expand_values(vals.shape[0], vals.shape[1], len(holder), _i32(holder), vals, out)
    return out


def dot(a, b, width=64):
    a = _f64(a)
    return lib().orc_dot(a.size, a, _f64(b), width)


def spmv(pat, vals, x, width=1):
    y = np.empty(pat.n)
    lib().orc_spmv(pat.n, pat.row_ptrs, pat.col_idx, _f64(vals), _f64(x), y, width)
    return y


def ilu0(pat, vals):
    F = _f64(vals).copy()
    st = lib().orc_ilu0(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, F)
    return F, st


def parilu0(pat, vals, sweeps, init=None):
    A = _f64(vals)
    F = (A if init is None else _f64(init)).copy()
    lib().orc_parilu0_sweeps(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, A, F, sweeps)
    return F


def factor_residual(pat, A, F):
    return lib().orc_factor_residual(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, _f64(A), _f64(F))


def lu_solve(pat, F, v):
    tmp = np.empty(pat.n)
    out = np.empty(pat.n)
    lib().orc_lu_solve(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, _f64(F), _f64(v), tmp, out)
    return out


def trsv_levels(pat):
    ll = np.empty(pat.n, np.int32)
    lu = np.empty(pat.n, np.int32)
    lib().orc_trsv_levels(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, ll, lu)
    return ll, lu


def lu_solve_avail(pat, F, v):
    ll, lu = trsv_levels(pat)
    tmp = np.empty(pat.n)
    out = np.empty(pat.n)
    lib().orc_lu_solve_avail(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, ll, lu, _f64(F), _f64(v),
                             tmp, out)
    return out


def isai(pat, vals, typ):
    out = np.zeros(pat.nnz)
    st = lib().orc_isai(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, _f64(vals), typ, out)
    return out, st


def bicgstab(pat, A, b, pc_kind=PC_NONE, P1=None, P2=None, o: Opts | None = None):
    o = o or opts()
    x = np.empty(pat.n)
    it, cv, ir, tr = C.c_float(), C.c_int(), C.c_double(), C.c_double()
    p1 = _f64(P1) if P1 is not None else None
    p2 = _f64(P2) if P2 is not None else None
    lib().orc_bicgstab(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, _f64(A), _f64(b), x, pc_kind,
                       p1.ctypes.data if p1 is not None else None,
                       p2.ctypes.data if p2 is not None else None,
                       C.byref(o), C.byref(it), C.byref(cv), C.byref(ir), C.byref(tr))
    return x, it.value, cv.value, ir.value, tr.value


def batch_solve(pat, vals, b, precond=ILU0, sweeps=5, o: Opts | None = None, threads=1):
    o = o or opts()
    vals = _f64(vals)
    b = _f64(b)
    B = b.shape[0]
    x = np.empty((B, pat.n))
    it = np.empty(B, dtype=np.float32)
    cv = np.empty(B, dtype=np.int32)
    ir = np.empty(B)
    tr = np.empty(B)
    st = lib().orc_batch_solve(pat.n, pat.nnz, pat.row_ptrs, pat.col_idx, B, vals, b, x, precond,
                               sweeps, C.byref(o), it, cv, ir, tr, threads)
    if st != 0:
        raise RuntimeError(f"orc_batch_solve failed: {st}")
    return x, it, cv, ir, tr


def entry_order(pat):
    """Entries with work of the persistent factor kernel, in its order."""
    order = np.empty(max(pat.nnz, 1), np.int32)
    ne = lib().orc_entry_order(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, order)
    return order[:ne]



def sweep_order(pat):
    """entries with work sorted by product count (fewest first), stable in
    the entry order: the factor kernel's ParILU round order"""
    order = np.empty(pat.nnz, np.int32)
    ne = lib().orc_sweep_order(pat.n, _i32(pat.row_ptrs), _i32(pat.col_idx), _i32(pat.diag), order)
    return order[:ne]

def parilu0_stop(pat, vals, sweeps, stop_tol=0.0, init=None, inplace=False):
    """ParILU(0) sweeps with warm start / residual stop / exact fallback,
    Jacobi or in place (rounds of the factor kernel's order, each reading
    what the earlier rounds wrote).  Returns (F, status): status = sweeps
    used or -1 (fallback)."""
    A = _f64(vals)
    F = np.empty(pat.nnz)
    st = C.c_int()
    ini = _f64(init) if init is not None else None
    lib().orc_parilu0_stop(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, A,
                           ini.ctypes.data if ini is not None else None, F, sweeps, stop_tol,
                           C.byref(st), int(bool(inplace)))
    return F, st.value


PREC_FP64, PREC_FP32_IR = 0, 1


def batch_solve_split(pat, vals, b, precond=ILU0, sweeps=5, init=None, stop_tol=0.0,
                      precision=PREC_FP64, inner_tol=1e-5, max_refine=10, o: Opts | None = None,
                      threads=1, inplace=False):
    """Split factor + solve paths (persistent factor kernel, then the FP64
    solve in availability order or the FP32-inner / FP64-refinement solve).
    Returns x, iter, conv, iter_res, true_res, refine, factor_status."""
    o = o or opts(trsv_order=TRSV_AVAIL)
    vals = _f64(vals)
    b = _f64(b)
    B = b.shape[0]
    x = np.empty((B, pat.n))
    it = np.empty(B, dtype=np.float32)
    cv = np.empty(B, dtype=np.int32)
    ir = np.empty(B)
    tr = np.empty(B)
    rf = np.empty(B, dtype=np.int32)
    fs = np.empty(B, dtype=np.int32)
    ini = _f64(init) if init is not None else None
    st = lib().orc_batch_solve_split(pat.n, pat.nnz, pat.row_ptrs, pat.col_idx, B, vals, b, x,
                                     {ILU0: 0, PARILU0: 1, ILU0_ISAI: 2, PARILU0_ISAI: 3, GENERAL_ISAI: 4}[precond], sweeps,
                                     ini.ctypes.data if ini is not None else None, stop_tol,
                                     precision, inner_tol, max_refine, C.byref(o), it, cv, ir, tr,
                                     rf, fs, threads, int(bool(inplace)))
    if st != 0:
        raise RuntimeError(f"orc_batch_solve_split failed: {st}")
    return x, it, cv, ir, tr, rf, fs


def isai_lu_f32(pat, F):
    """FP32 ISAI(L) / ISAI(U) of the in-place factor F on one array; returns
    (M, status)."""
    M = np.empty(pat.nnz, np.float32)
    st = lib().orc_isai_lu_f32(pat.n, pat.row_ptrs, pat.col_idx, pat.diag, _f64(F), M)
    return M, st
