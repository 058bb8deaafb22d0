"""Python binding of libbilu (ctypes over the C ABI in include/bilu.h).

Device buffers are torch tensors on a HIP device (torch is used only for
device memory and streams).  The HIP library is required: if libbilu.so is
missing or cannot be loaded, every entry point raises -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbilu.so")

OK = 0
ERR_INVALID_ARG, ERR_INVALID_PATTERN, ERR_MISSING_DIAGONAL, ERR_TOO_LARGE, ERR_HIP, ERR_NO_DEVICE = 1, 2, 3, 4, 5, 6
CONVERGED, NOT_CONVERGED, ZERO_PIVOT, NONFINITE = 1, -1, -2, -3
ILU0, PARILU0, ILU0_ISAI, PARILU0_ISAI, GENERAL_ISAI, NONE, FACTORS = 0, 1, 2, 3, 4, 5, 6
TOL_ABSOLUTE, TOL_RELATIVE = 0, 1
MERGED, SEGREGATED, MERGE_AUTO = 0, 1, 2
ISAI_GENERAL, ISAI_LOWER, ISAI_UPPER = 0, 1, 2
KERNEL_AUTO, KERNEL_WAVE, KERNEL_PERSISTENT = 0, 1, 2
PRECISION_FP64, PRECISION_FP32_IR = 0, 1
PARILU_JACOBI, PARILU_INPLACE = 0, 1
TIMING_FACTOR, TIMING_SOLVE = 1, 2
ILU_SCHED_AUTO, ILU_SCHED_ROWS, ILU_SCHED_ENTRIES, ILU_SCHED_ERAS, ILU_SCHED_STREAMS = 0, 1, 2, 3, 4

# every symbol include/bilu.h declares
EXPORTS = [
    "bilu_opts_default", "bilu_version", "bilu_status_string", "bilu_pattern_create",
    "bilu_pattern_destroy", "bilu_pattern_get_info", "bilu_pattern_lu_split", "bilu_solve",
    "bilu_solve_ex", "bilu_ilu0", "bilu_parilu0", "bilu_factor_residual", "bilu_split_lu", "bilu_isai",
    "bilu_true_residual", "bilu_mlp_create", "bilu_mlp_destroy", "bilu_mlp_num_params",
    "bilu_mlp_get_weights", "bilu_mlp_set_weights", "bilu_mlp_gradient", "bilu_mlp_train_step",
    "bilu_mlp_infer", "bilu_uneven_create", "bilu_uneven_destroy", "bilu_uneven_tune", "bilu_uneven_config",
    "bilu_uneven_merge_point", "bilu_uneven_placement", "bilu_uneven_tune_table", "bilu_uneven_configure",
    "bilu_uneven_solve", "bilu_pattern_add_diagonal", "bilu_expand_values", "bilu_isai_pattern",
    "bilu_isai_power", "bilu_mlp_solve", "bilu_ilu0_ex", "bilu_ilu0_tune", "bilu_pattern_reserve",
    "bilu_solve_ex_noalloc", "bilu_timing_begin", "bilu_timing_end",
]


class BiluError(RuntimeError):
    pass


class Opts(C.Structure):
    _fields_ = [
        ("precond", C.c_int),
        ("parilu_sweeps", C.c_int),
        ("tol", C.c_double),
        ("tol_mode", C.c_int),
        ("max_iter", C.c_int),
        ("merge", C.c_int),
        ("true_res_restart", C.c_int),
        ("max_restarts", C.c_int),
        ("kernel", C.c_int),
        ("precision", C.c_int),
        ("inner_tol", C.c_double),
        ("max_refine", C.c_int),
        ("parilu_stop_tol", C.c_double),
        ("isai_power", C.c_int),
        ("inner_adapt", C.c_double),
        ("parilu_order", C.c_int),
    ]


class PatternInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "nnz", "l_nnz", "u_nnz", "levels_lower",
                                        "levels_upper", "ilu_updates", "max_row_nnz",
                                        "levels_entries", "ilu_eras")]


_lib = None


def lib():
    """Load libbilu.so (raises if it is missing: no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BiluError(f"libbilu.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32p = C.c_void_p, C.POINTER(C.c_int)
        L.bilu_opts_default.argtypes = [C.POINTER(Opts)]
        L.bilu_opts_default.restype = None
        L.bilu_version.restype = C.c_char_p
        L.bilu_status_string.argtypes = [C.c_int]
        L.bilu_status_string.restype = C.c_char_p
        L.bilu_pattern_create.argtypes = [C.c_int, C.c_int, i32p, i32p, C.POINTER(vp)]
        L.bilu_pattern_destroy.argtypes = [vp]
        L.bilu_pattern_get_info.argtypes = [vp, C.POINTER(PatternInfo)]
        L.bilu_pattern_reserve.argtypes = [vp, C.c_int]
        L.bilu_pattern_lu_split.argtypes = [vp, i32p, i32p, i32p, i32p]
        L.bilu_pattern_add_diagonal.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p, i32p, i32p]
        L.bilu_expand_values.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.bilu_solve.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(Opts), vp, vp, vp, vp, vp]
        L.bilu_solve_ex.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(Opts), vp, vp, vp, vp, vp,
                                    vp, vp, vp]
        L.bilu_solve_ex_noalloc.argtypes = L.bilu_solve_ex.argtypes
        L.bilu_timing_begin.argtypes = [C.c_int, C.c_int]
        L.bilu_timing_end.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.bilu_ilu0.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.bilu_parilu0.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, vp]
        L.bilu_factor_residual.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.bilu_split_lu.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.bilu_isai.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp, vp]
        L.bilu_isai_pattern.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), i32p, i32p]
        L.bilu_isai_power.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp]
        L.bilu_true_residual.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.bilu_uneven_create.argtypes = [C.c_int, C.POINTER(vp), C.c_int, i32p, i32p, C.POINTER(Opts),
                                         C.POINTER(vp)]
        L.bilu_uneven_destroy.argtypes = [vp]
        L.bilu_uneven_tune.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp, vp, vp, vp, vp, vp]
        L.bilu_uneven_config.argtypes = [vp, C.c_int, i32p, i32p, i32p]
        L.bilu_uneven_merge_point.argtypes = [vp, C.c_int, i32p]
        L.bilu_uneven_placement.argtypes = [vp, C.c_int, i32p]
        L.bilu_uneven_configure.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.bilu_uneven_tune_table.argtypes = [vp, C.c_int, C.c_int, i32p, i32p, i32p, i32p, C.POINTER(C.c_float),
                                             C.POINTER(C.c_float), i32p]
        L.bilu_uneven_solve.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp, vp, vp, vp, vp, vp]
        L.bilu_mlp_create.argtypes = [vp, C.c_int, C.c_ulonglong, C.POINTER(vp)]
        L.bilu_mlp_destroy.argtypes = [vp]
        L.bilu_mlp_num_params.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.bilu_mlp_get_weights.argtypes = [vp, vp, vp]
        L.bilu_mlp_set_weights.argtypes = [vp, vp, vp]
        L.bilu_mlp_gradient.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.bilu_mlp_train_step.argtypes = [vp, C.c_int, vp, vp, C.c_float, vp, vp]
        L.bilu_mlp_infer.argtypes = [vp, C.c_int, vp, vp, vp]
        L.bilu_ilu0_ex.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, vp]
        L.bilu_ilu0_tune.argtypes = [vp, C.c_int, vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_int), vp]
        L.bilu_mlp_solve.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.POINTER(Opts), vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != OK:
        msg = lib().bilu_status_string(rc).decode()
        raise BiluError(f"{what} failed: {rc} ({msg})")


def opts(precond=ILU0, parilu_sweeps=5, tol=1e-11, tol_mode=TOL_ABSOLUTE, max_iter=500,
         merge=MERGE_AUTO, true_res_restart=0, max_restarts=3, kernel=KERNEL_AUTO,
         precision=PRECISION_FP64, inner_tol=1e-5, max_refine=10, parilu_stop_tol=0.0,
         isai_power=1, parilu_order=PARILU_JACOBI, inner_adapt=0.5) -> Opts:
    return Opts(precond, parilu_sweeps, tol, tol_mode, max_iter, merge, true_res_restart,
                max_restarts, kernel, precision, inner_tol, max_refine, parilu_stop_tol, isai_power,
                inner_adapt, parilu_order)


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Pattern:
    """Handle of an analysed shared CSR pattern on the current device."""

    def __init__(self, n, row_ptrs, col_idx):
        rp = np.ascontiguousarray(row_ptrs, dtype=np.int32)
        ci = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.n = int(n)
        self.nnz = int(rp[-1])
        self.row_ptrs, self.col_idx = rp, ci
        h = C.c_void_p()
        _check(lib().bilu_pattern_create(self.n, self.nnz, rp.ctypes.data_as(C.POINTER(C.c_int)),
                                         ci.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h)),
               "bilu_pattern_create")
        self.h = h

    @classmethod
    def from_synth(cls, pat):
        return cls(pat.n, pat.row_ptrs, pat.col_idx)

    def info(self) -> PatternInfo:
        inf = PatternInfo()
        _check(lib().bilu_pattern_get_info(self.h, C.byref(inf)), "bilu_pattern_get_info")
        return inf

    def reserve(self, batch):
        """size the solve workspace for batches of up to `batch` systems (no
        allocation or device wait inside later solves, e.g. under graph capture)"""
        _check(lib().bilu_pattern_reserve(self.h, int(batch)), "bilu_pattern_reserve")

    def lu_split(self):
        inf = self.info()
        lrp = np.empty(self.n + 1, np.int32)
        lci = np.empty(inf.l_nnz, np.int32)
        urp = np.empty(self.n + 1, np.int32)
        uci = np.empty(inf.u_nnz, np.int32)
        P = C.POINTER(C.c_int)
        _check(lib().bilu_pattern_lu_split(self.h, lrp.ctypes.data_as(P), lci.ctypes.data_as(P),
                                           urp.ctypes.data_as(P), uci.ctypes.data_as(P)),
               "bilu_pattern_lu_split")
        return lrp, lci, urp, uci

    def close(self):
        if getattr(self, "h", None):
            lib().bilu_pattern_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def add_diagonal(n, row_ptrs, col_idx):
    """Pattern with every missing diagonal entry inserted (host, no device
    needed): returns (row_ptrs, col_idx, src), src[k] = position in the input
    pattern or -1 for an inserted (i,i)."""
    rp = np.ascontiguousarray(row_ptrs, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    P = C.POINTER(C.c_int)
    miss = C.c_int()
    _check(lib().bilu_pattern_add_diagonal(n, len(ci), rp.ctypes.data_as(P), ci.ctypes.data_as(P), C.byref(miss),
                                           None, None, None), "bilu_pattern_add_diagonal")
    rpo = np.empty(n + 1, np.int32)
    cio = np.empty(len(ci) + miss.value, np.int32)
    src = np.empty(len(ci) + miss.value, np.int32)
    _check(lib().bilu_pattern_add_diagonal(n, len(ci), rp.ctypes.data_as(P), ci.ctypes.data_as(P), C.byref(miss),
                                           rpo.ctypes.data_as(P), cio.ctypes.data_as(P), src.ctypes.data_as(P)),
           "bilu_pattern_add_diagonal")
    return rpo, cio, src


def expand_values(src, vals, out, stream=None):
    """out[s, k] = 0 if src[k] < 0 else vals[s, src[k]] (device tensors)"""
    _check(lib().bilu_expand_values(vals.shape[0], vals.shape[1], out.shape[1], _ptr(src), _ptr(vals), _ptr(out),
                                    _stream(stream)), "bilu_expand_values")


def solve(pat: Pattern, A, b, x, o: Opts | None = None, iters=None, conv=None, iter_res=None,
          true_res=None, stream=None):
    """Batched solve on device tensors; returns (iters, conv, iter_res, true_res)."""
    import torch
    B = b.shape[0]
    dev = b.device
    if iters is None:
        iters = torch.empty(B, dtype=torch.float32, device=dev)
    if conv is None:
        conv = torch.empty(B, dtype=torch.int32, device=dev)
    if iter_res is None:
        iter_res = torch.empty(B, dtype=torch.float64, device=dev)
    if true_res is None:
        true_res = torch.empty(B, dtype=torch.float64, device=dev)
    o = o or opts()
    _check(lib().bilu_solve(pat.h, B, _ptr(A), _ptr(b), _ptr(x), C.byref(o), _ptr(iters),
                            _ptr(conv), _ptr(iter_res), _ptr(true_res), _stream(stream)),
           "bilu_solve")
    return iters, conv, iter_res, true_res


def solve_ex(pat: Pattern, A, b, x, o: Opts | None = None, F_init=None, iters=None, conv=None,
             iter_res=None, true_res=None, refine=None, factor_status=None, stream=None):
    """bilu_solve_ex: mixed precision / warm-started ParILU.  Returns
    (iters, conv, iter_res, true_res, refine, factor_status)."""
    import torch
    B = b.shape[0]
    dev = b.device
    mk = lambda t, dt: t if t is not None else torch.empty(B, dtype=dt, device=dev)
    iters, conv = mk(iters, torch.float32), mk(conv, torch.int32)
    iter_res, true_res = mk(iter_res, torch.float64), mk(true_res, torch.float64)
    refine = mk(refine, torch.int32)
    # factor_status=False: not requested (a plain solve then keeps its merged kernel)
    factor_status = None if factor_status is False else mk(factor_status, torch.int32)
    o = o or opts()
    _check(lib().bilu_solve_ex(pat.h, B, _ptr(A), _ptr(b), _ptr(x), C.byref(o), _ptr(F_init),
                               _ptr(iters), _ptr(conv), _ptr(iter_res), _ptr(true_res),
                               _ptr(refine), _ptr(factor_status), _stream(stream)),
           "bilu_solve_ex")
    return iters, conv, iter_res, true_res, refine, factor_status


def solve_ex_noalloc(pat: Pattern, A, b, x, o: Opts | None = None, F_init=None, iters=None, conv=None,
                     iter_res=None, true_res=None, refine=None, factor_status=None, stream=None):
    """bilu_solve_ex_noalloc: bilu_solve_ex without any allocation (the
    library's fixed per-GPU scratch array).  Outputs as solve_ex; pass every
    output tensor to keep the call itself free of torch allocations."""
    import torch
    B = b.shape[0]
    dev = b.device
    mk = lambda t, dt: t if t is not None else torch.empty(B, dtype=dt, device=dev)
    iters, conv = mk(iters, torch.float32), mk(conv, torch.int32)
    iter_res, true_res = mk(iter_res, torch.float64), mk(true_res, torch.float64)
    refine = mk(refine, torch.int32)
    factor_status = None if factor_status is False else mk(factor_status, torch.int32)
    o = o or opts()
    _check(lib().bilu_solve_ex_noalloc(pat.h, B, _ptr(A), _ptr(b), _ptr(x), C.byref(o), _ptr(F_init),
                                       _ptr(iters), _ptr(conv), _ptr(iter_res), _ptr(true_res),
                                       _ptr(refine), _ptr(factor_status), _stream(stream)),
           "bilu_solve_ex_noalloc")
    return iters, conv, iter_res, true_res, refine, factor_status


class KernelTiming:
    """with KernelTiming(TIMING_SOLVE, n) as t: ... -> t.avg_ms, t.launches: the
    average duration of that kernel's launches inside the block, from HIP
    events the library records on the kernel's stream"""

    def __init__(self, kernel, max_launches):
        self.kernel, self.max_launches = kernel, max_launches
        self.avg_ms, self.launches = None, 0

    def __enter__(self):
        _check(lib().bilu_timing_begin(self.kernel, self.max_launches), "bilu_timing_begin")
        return self

    def __exit__(self, *exc):
        ms, n = C.c_float(), C.c_int()
        _check(lib().bilu_timing_end(C.byref(ms), C.byref(n)), "bilu_timing_end")
        self.avg_ms, self.launches = ms.value, n.value
        return False


def ilu0(pat: Pattern, A, F, status=None, stream=None, schedule=None):
    """exact ILU(0) in place layout; schedule = ILU_SCHED_* (None: bilu_ilu0,
    the pattern's tuned schedule)"""
    if schedule is None:
        _check(lib().bilu_ilu0(pat.h, A.shape[0], _ptr(A), _ptr(F), _ptr(status), _stream(stream)),
               "bilu_ilu0")
    else:
        _check(lib().bilu_ilu0_ex(pat.h, A.shape[0], _ptr(A), _ptr(F), schedule, _ptr(status),
                                  _stream(stream)), "bilu_ilu0_ex")


def ilu0_tune(pat: Pattern, A, F, stream=None):
    """bilu_ilu0_tune: returns (best schedule, [ms per run of ROWS, ENTRIES,
    ERAS, STREAMS], -1 = not supported)"""
    ms = (C.c_float * 4)()
    best = C.c_int()
    _check(lib().bilu_ilu0_tune(pat.h, A.shape[0], _ptr(A), _ptr(F), ms, C.byref(best), _stream(stream)),
           "bilu_ilu0_tune")
    return best.value, list(ms)


def parilu0(pat: Pattern, A, F, sweeps, F_init=None, stream=None):
    _check(lib().bilu_parilu0(pat.h, A.shape[0], _ptr(A), _ptr(F_init), _ptr(F), sweeps,
                              _stream(stream)), "bilu_parilu0")


def factor_residual(pat: Pattern, A, F, res, stream=None):
    _check(lib().bilu_factor_residual(pat.h, A.shape[0], _ptr(A), _ptr(F), _ptr(res),
                                      _stream(stream)), "bilu_factor_residual")


def split_lu(pat: Pattern, F, L, U, stream=None):
    _check(lib().bilu_split_lu(pat.h, F.shape[0], _ptr(F), _ptr(L), _ptr(U), _stream(stream)),
           "bilu_split_lu")


def isai(pat: Pattern, vals, typ, out, status=None, stream=None):
    _check(lib().bilu_isai(pat.h, vals.shape[0], _ptr(vals), typ, _ptr(out), _ptr(status),
                           _stream(stream)), "bilu_isai")


def isai_pattern(pat: Pattern, typ, power):
    """host (row_ptrs, col_idx) of the ISAI pattern |P|^power of the typed view"""
    P = C.POINTER(C.c_int)
    nnz = C.c_int()
    _check(lib().bilu_isai_pattern(pat.h, typ, power, C.byref(nnz), None, None), "bilu_isai_pattern")
    rp, ci = np.empty(pat.n + 1, np.int32), np.empty(nnz.value, np.int32)
    _check(lib().bilu_isai_pattern(pat.h, typ, power, C.byref(nnz), rp.ctypes.data_as(P), ci.ctypes.data_as(P)),
           "bilu_isai_pattern")
    return rp, ci


def isai_power(pat: Pattern, vals, typ, power, out, status=None, stream=None):
    """ISAI of the typed view of vals [B, nnz] on |P|^power: out [B, nnz_power]"""
    _check(lib().bilu_isai_power(pat.h, vals.shape[0], _ptr(vals), typ, power, _ptr(out), _ptr(status),
                                 _stream(stream)), "bilu_isai_power")


def true_residual(pat: Pattern, A, b, x, res, stream=None):
    _check(lib().bilu_true_residual(pat.h, A.shape[0], _ptr(A), _ptr(b), _ptr(x), _ptr(res),
                                    _stream(stream)), "bilu_true_residual")


class Mlp:
    """Learned warm start of ParILU(0) (bilu_mlp_*): predicts in-place ILU(0)
    factors from A; trained on exact factors with the library's BF16 MFMA
    kernels."""

    def __init__(self, pat: Pattern, hidden=512, seed=1):
        h = C.c_void_p()
        _check(lib().bilu_mlp_create(pat.h, hidden, seed, C.byref(h)), "bilu_mlp_create")
        self.h = h
        self.nnz = pat.nnz
        self.hidden = hidden
        n = C.c_longlong()
        _check(lib().bilu_mlp_num_params(self.h, C.byref(n)), "bilu_mlp_num_params")
        self.num_params = int(n.value)

    def close(self):
        if getattr(self, "h", None):
            lib().bilu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self, stream=None):
        import torch
        w = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        _check(lib().bilu_mlp_get_weights(self.h, _ptr(w), _stream(stream)), "bilu_mlp_get_weights")
        return w

    def set_weights(self, w, stream=None):
        _check(lib().bilu_mlp_set_weights(self.h, _ptr(w), _stream(stream)), "bilu_mlp_set_weights")

    def gradient(self, A, F, stream=None):
        import torch
        loss = torch.empty(1, dtype=torch.float32, device=A.device)
        g = torch.empty(self.num_params, dtype=torch.float32, device=A.device)
        _check(lib().bilu_mlp_gradient(self.h, A.shape[0], _ptr(A), _ptr(F), _ptr(loss), _ptr(g),
                                       _stream(stream)), "bilu_mlp_gradient")
        return loss, g

    def train_step(self, A, F, lr, loss=None, stream=None):
        _check(lib().bilu_mlp_train_step(self.h, A.shape[0], _ptr(A), _ptr(F), float(lr), _ptr(loss),
                                         _stream(stream)), "bilu_mlp_train_step")

    def infer(self, A, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(A)
        _check(lib().bilu_mlp_infer(self.h, A.shape[0], _ptr(A), _ptr(out), _stream(stream)),
               "bilu_mlp_infer")
        return out

    def solve(self, pat: Pattern, A, b, x, o: Opts, iters=None, conv=None, iter_res=None, true_res=None,
              refine=None, factor_status=None, stream=None):
        """bilu_mlp_solve: infer + FP32-IR ParILU(0) solve in one call, the same
        results as infer() followed by solve_ex(F_init=...).  Returns
        (iters, conv, iter_res, true_res, refine, factor_status)."""
        import torch
        B = b.shape[0]
        dev = b.device
        mk = lambda t, dt: t if t is not None else torch.empty(B, dtype=dt, device=dev)
        iters, conv = mk(iters, torch.float32), mk(conv, torch.int32)
        iter_res, true_res = mk(iter_res, torch.float64), mk(true_res, torch.float64)
        refine = mk(refine, torch.int32)
        factor_status = None if factor_status is False else mk(factor_status, torch.int32)
        _check(lib().bilu_mlp_solve(self.h, pat.h, B, _ptr(A), _ptr(b), _ptr(x), C.byref(o), _ptr(iters),
                                    _ptr(conv), _ptr(iter_res), _ptr(true_res), _ptr(refine),
                                    _ptr(factor_status), _stream(stream)), "bilu_mlp_solve")
        return iters, conv, iter_res, true_res, refine, factor_status


class Uneven:
    """Plan of an uneven request list over several patterns (bilu_uneven_*):
    request r solves system req_index[r] of pattern req_pattern[r]."""

    def __init__(self, patterns, req_pattern, req_index, o: Opts):
        self.patterns = list(patterns)
        hs = (C.c_void_p * len(self.patterns))(*[p.h.value for p in self.patterns])
        rp = np.ascontiguousarray(req_pattern, dtype=np.int32)
        ri = np.ascontiguousarray(req_index, dtype=np.int32)
        self.nreq = int(rp.size)
        h = C.c_void_p()
        _check(lib().bilu_uneven_create(len(self.patterns), hs, self.nreq,
                                        rp.ctypes.data_as(C.POINTER(C.c_int)),
                                        ri.ctypes.data_as(C.POINTER(C.c_int)), C.byref(o), C.byref(h)),
               "bilu_uneven_create")
        self.h = h

    def _ptrs(self, ts):
        return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def _call(self, fn, A, b, x, out, stream):
        it, cv, ir, tr, rf = out
        _check(fn(self.h, self._ptrs(A), self._ptrs(b), self._ptrs(x), _ptr(it), _ptr(cv), _ptr(ir),
                  _ptr(tr), _ptr(rf), _stream(stream)), fn.__name__)

    def outputs(self, device="cuda"):
        import torch
        B = self.nreq
        return (torch.empty(B, dtype=torch.float32, device=device), torch.empty(B, dtype=torch.int32, device=device),
                torch.empty(B, dtype=torch.float64, device=device), torch.empty(B, dtype=torch.float64, device=device),
                torch.empty(B, dtype=torch.int32, device=device))

    def tune(self, A, b, x, out, stream=None):
        self._call(lib().bilu_uneven_tune, A, b, x, out, stream)

    def solve(self, A, b, x, out, stream=None):
        self._call(lib().bilu_uneven_solve, A, b, x, out, stream)

    def config(self, p):
        """(waves per factor workgroup, 0 for the merged kernel; waves per solve
        workgroup; requests) of pattern p"""
        fw, sw, jobs = C.c_int(), C.c_int(), C.c_int()
        _check(lib().bilu_uneven_config(self.h, p, C.byref(fw), C.byref(sw), C.byref(jobs)), "bilu_uneven_config")
        return fw.value, sw.value, jobs.value

    def merged(self, p):
        m = C.c_int()
        _check(lib().bilu_uneven_merge_point(self.h, p, C.byref(m)), "bilu_uneven_merge_point")
        return bool(m.value)

    def configure(self, p, merged, factor_waves, solve_waves, records_global=False):
        _check(lib().bilu_uneven_configure(self.h, p, int(merged), factor_waves, solve_waves, int(records_global)),
               "bilu_uneven_configure")

    def records_global(self, p):
        g = C.c_int()
        _check(lib().bilu_uneven_placement(self.h, p, C.byref(g)), "bilu_uneven_placement")
        return bool(g.value)

    def tune_table(self, p):
        """the candidates the last tune() timed for pattern p: dicts of merged,
        factor_waves, solve_waves, records_global, alone_ms, corun_ms"""
        cnt = C.c_int()
        _check(lib().bilu_uneven_tune_table(self.h, p, 0, None, None, None, None, None, None, C.byref(cnt)),
               "bilu_uneven_tune_table")
        k = cnt.value
        mg, fw, sw, rg = ((C.c_int * max(k, 1))() for _ in range(4))
        al, co = (C.c_float * max(k, 1))(), (C.c_float * max(k, 1))()
        _check(lib().bilu_uneven_tune_table(self.h, p, k, mg, fw, sw, rg, al, co, C.byref(cnt)),
               "bilu_uneven_tune_table")
        return [{"merged": bool(mg[i]), "factor_waves": fw[i], "solve_waves": sw[i], "records_global": bool(rg[i]),
                 "alone_ms": round(al[i], 4), "corun_ms": round(co[i], 4)} for i in range(k)]

    def close(self):
        if getattr(self, "h", None):
            lib().bilu_uneven_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
