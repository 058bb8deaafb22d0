"""Benchmark: systems solved per second to relres <= 1e-10 on a batch of
65536 sparse systems (n=144, one shared random pattern, ~2000 nnz), one GPU.

One step = one batched solve over the whole batch, inputs (A, b) resident
in HBM: preconditioner setup (and, on the learned path, the network's
inference) + BiCGSTAB + FP64 true residual.  Prints one JSON line (see the
repository's bench contract).

Default (headline): the learned + mixed-precision solve -- MLP warm start,
in-place ParILU(0) sweeps to a factor residual of 1e-8 (exact ILU(0)
fallback), ISAI(L, U) of the factors applied in FP32, FP32 BiCGSTAB inside
FP64 iterative refinement -- timed beside the plain FP64 merged exact-ILU(0)
BiCGSTAB kernel and the mixed-precision solve without the network.  The
network is trained before the timed region (reported under "training", with
held-out sweep counts against its untrained base guess and against no warm
start).  The line also carries the uneven-batch result (4 patterns), the
FP32-IR solution error against dense FP64 solves, the dominant kernel's
duration from HIP events the library records on its stream, and the CPU
baseline (the oracle on every core of this host).
--mode single times one variant chosen by the flags; --mode uneven only the
uneven batch.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "merging-ilu0_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth
METRIC = "systems solved/s at relres<=1e-10, n=144 batch 65536: learned+mixed vs plain FP64"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--n", type=int, default=144)
    ap.add_argument("--mode", default="headline", choices=["headline", "single", "uneven", "learned"],
                    help="learned: only the learned + mixed step (training as --train-*), for profiling")
    ap.add_argument("--uneven-each", type=int, default=16384)
    ap.add_argument("--uneven-precond", default="ilu0_isai", choices=["ilu0", "ilu0_isai"],
                    help="preconditioner of the uneven leg (ISAI(L, U) applies or triangular solves)")
    ap.add_argument("--precond", default="ilu0",
                    choices=["ilu0", "parilu0", "ilu0_isai", "parilu0_isai", "general_isai"])
    ap.add_argument("--merge", default="auto", choices=["auto", "merged", "segregated"],
                    help="--mode single: the merge point (auto = the library's default rule)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "wave", "persistent"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir"],
                    help="fp32_ir: FP32 preconditioner and inner BiCGSTAB inside FP64 refinement")
    ap.add_argument("--inner-tol", type=float, default=1e-5)
    ap.add_argument("--isai-power", type=int, default=1,
                    help="--mode single, ISAI preconditioners: sparsity of |L|^k, |U|^k (|A|^k)")
    ap.add_argument("--train-systems", type=int, default=262144)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a 1 GiB device copy (HBM counter calibration)")
    return ap.parse_args()


def committed_profile(key):
    """Counter figures of a workload from the committed rocprofv3 --pmc runs
    (profiles/counters.json: HBM traffic per launch from FETCH_SIZE /
    WRITE_SIZE scaled by a device copy's reading, LDS bytes per launch from
    the SQ_INSTS_LDS_*_BANDWIDTH counters).  These are not measured by this
    run; the record says where they come from.  None when not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "counters.json")) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def traffic_fields(key):
    rec = committed_profile(key)
    if rec is None:
        return {"traffic": None, "traffic_source": None}
    return {"traffic": rec.get("traffic_bytes_per_launch"),
            "traffic_source": f"profiles/counters.json[{key!r}] (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                              f"committed, not measured by this run)"}


# LDS peak of MI355X: 256 CUs x 128 bytes per clock x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4


def lds_fields(key, kernel_ms):
    """the kernel's LDS bytes per launch (committed counters) over its
    duration measured in this run, against the chip's LDS peak"""
    rec = committed_profile(key)
    if rec is None or kernel_ms is None or "lds_bytes_per_launch" not in rec:
        return None
    gbs = rec["lds_bytes_per_launch"] / (kernel_ms * 1e-3) / 1e9
    return {"achieved": round(gbs, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / LDS_PEAK_GBS, 4),
            "lds_bytes_per_launch": rec["lds_bytes_per_launch"],
            "source": f"LDS bytes: profiles/counters.json[{key!r}] (committed); duration: this run"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(pat, A, b, precond):
    """The oracle (CPU restatement of the reference path, same algorithm and
    stopping rule) timed on this host: the whole batch on every core this
    process may use (a few seconds of CPU work), plus a single-core figure
    on a prefix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    code = {"ilu0": O.ILU0, "parilu0": O.PARILU0, "ilu0_isai": O.ILU0_ISAI,
            "parilu0_isai": O.PARILU0_ISAI, "general_isai": O.GENERAL_ISAI}[precond]
    oo = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1)
    threads = host_cores()
    s_all = A.shape[0]
    t = time.perf_counter()
    O.batch_solve(pat, A[:s_all], b[:s_all], code, sweeps=5, o=oo, threads=threads)
    dt_all = time.perf_counter() - t
    s_one = min(A.shape[0], 16384)
    t = time.perf_counter()
    O.batch_solve(pat, A[:s_one], b[:s_one], code, sweeps=5, o=oo, threads=1)
    dt_one = time.perf_counter() - t
    return {"value": round(s_all / dt_all, 1), "unit": "systems/s", "cores": threads,
            "kind": "port",
            "sample": f"{s_all} systems (the whole batch), plain FP64 {precond} + BiCGSTAB, "
                      f"{threads} threads (every core of the affinity mask), {dt_all:.2f} s wall",
            "single_core_value": round(s_one / dt_one, 1),
            "single_core_sample": f"first {s_one} systems, 1 thread, {dt_one:.2f} s",
            "cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count()}


def hbm_copy_peak(dev):
    """device-to-device copy of 2 GiB, timed with HIP events: the HBM rate a
    plain streaming kernel reaches on this GPU (read + write bytes / time)"""
    import torch
    a = torch.empty(1 << 28, dtype=torch.float64, device=dev)
    c = torch.empty_like(a)
    a.fill_(1.0)
    c.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        c.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    gbs = 2 * a.numel() * 8 / (ms * 1e-3) / 1e9
    del a, c
    torch.cuda.empty_cache()
    return round(gbs, 1)


def time_steps(step, warmup, steps, stream, kernel=None):
    """HIP events on the launch stream around exactly `steps` steps; with
    `kernel` (bilu.TIMING_*) also the average duration of that kernel's
    launches inside the timed region (events the library records on the
    kernel's stream).  Returns ms per step, or (ms per step, kernel ms)."""
    import torch
    import bilu
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    kt = bilu.KernelTiming(kernel, 4 * steps + 4) if kernel else None
    if kt:
        kt.__enter__()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    if kt:
        kt.__exit__(None, None, None)
    ms = e0.elapsed_time(e1) / max(steps, 1)
    return (ms, kt.avg_ms) if kt else ms


def uneven_bench(args, dev, precond=None):
    """Uneven batch: 4 patterns (n = 32, 90, 144, 300, nnz ~ 14 n), `uneven_each`
    systems each, interleaved in one request list, exact ILU(0) with its
    ISAI(L, U) applies (or its triangular solves, precond="ilu0") in FP32,
    FP32 inner / FP64 refinement; the plan is autotuned, then its hipGraph
    is replayed."""
    import numpy as np
    import torch

    import bilu
    import synth
    ns = (32, 90, 144, 300)
    E = args.uneven_each
    pats, dA, db, dx, nnzs = [], [], [], [], []
    for n in ns:
        pat, A, b = synth.batch("random", E, n=n, density=13.0 / (n - 1))
        pats.append(bilu.Pattern.from_synth(pat))
        dA.append(torch.from_numpy(A).to(dev)); db.append(torch.from_numpy(b).to(dev))
        dx.append(torch.empty_like(db[-1]))
        nnzs.append(pat.nnz)
    req_p = np.repeat(np.arange(len(ns)), E).astype(np.int32)
    req_i = np.tile(np.arange(E), len(ns)).astype(np.int32)
    perm = np.random.default_rng(0).permutation(req_p.size)
    req_p, req_i = req_p[perm], req_i[perm]
    precond = precond or args.uneven_precond
    pc = {"ilu0": bilu.ILU0, "ilu0_isai": bilu.ILU0_ISAI}[precond]
    plan = bilu.Uneven(pats, req_p, req_i, bilu.opts(precond=pc, tol=1e-10, tol_mode=bilu.TOL_RELATIVE,
                                                     precision=bilu.PRECISION_FP32_IR, inner_tol=args.inner_tol))
    out = plan.outputs()
    t0 = time.time()
    plan.tune(dA, db, dx, out)
    torch.cuda.synchronize()
    tune_s = time.time() - t0
    stream = torch.cuda.current_stream()
    ms = time_steps(lambda: plan.solve(dA, db, dx, out, stream=stream), args.warmup, args.steps, stream)
    cv, tr = out[1].cpu().numpy(), out[3].cpu().numpy()
    bn = np.concatenate([torch.linalg.norm(b, dim=1).cpu().numpy() for b in db])
    # per-request relative residual: request r is system req_i[r] of pattern req_p[r]
    offs = np.cumsum([0] + [E] * len(ns))
    relres = tr / bn[offs[req_p] + req_i]
    B = req_p.size
    alg = sum(E * (8 * nz + 16 * n + 24) for n, nz in zip(ns, nnzs)) + 4 * B
    achieved = alg / (ms * 1e-3) / 1e9
    return {"metric": "systems solved/s, uneven batch (4 patterns, n = 32/90/144/300), relres<=1e-10",
            "value": round(B / (ms * 1e-3), 1), "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"{len(ns)} patterns x {E} systems interleaved, exact ILU(0) "
                                   f"{'with ISAI(L, U) applies' if precond == 'ilu0_isai' else 'triangular solves'} "
                                   f"in FP32 + FP32 BiCGSTAB in FP64 refinement, one hipGraph",
                       "precond": precond, "n": list(ns), "nnz": nnzs},
            "all_converged_relres_le_1e-10": bool((cv == 1).all() and (relres <= 1e-10).all()),
            "tuned": {str(n): dict(zip(("factor_waves", "solve_waves", "jobs"), plan.config(p)),
                                   merged=plan.merged(p)) for p, n in enumerate(ns)},
            "tune_table": {str(n): plan.tune_table(p) for p, n in enumerate(ns)},
            "tune_seconds": round(tune_s, 2),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                         "kernel": "whole graph (per-pattern k_factor_fast + k_mixed)", "alg_bytes_per_launch": alg}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the batch shards "
              "into independent per-GPU slices with no exchange, so N GPUs = N independent runs.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import bilu
    import synth

    dev = torch.device("cuda:0")
    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        c = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": a.numel() * 8}))
        return

    if args.mode == "uneven":
        print(json.dumps(uneven_bench(args, dev)), flush=True)
        return

    t0 = time.time()
    pat, A, b = synth.batch("random", args.batch, n=args.n)
    gen_s = time.time() - t0
    P = bilu.Pattern.from_synth(pat)
    info = P.info()
    dA = torch.from_numpy(A).to(dev)
    db = torch.from_numpy(b).to(dev)
    x = torch.empty_like(db)
    B, n, nnz = args.batch, pat.n, pat.nnz
    bnorm = np.linalg.norm(b, axis=1)
    stream = torch.cuda.current_stream()
    out = dict(it=torch.empty(B, dtype=torch.float32, device=dev),
               cv=torch.empty(B, dtype=torch.int32, device=dev),
               ir=torch.empty(B, dtype=torch.float64, device=dev),
               tr=torch.empty(B, dtype=torch.float64, device=dev),
               rf=torch.zeros(B, dtype=torch.int32, device=dev),
               fs=torch.zeros(B, dtype=torch.int32, device=dev))
    PC = {"ilu0": bilu.ILU0, "parilu0": bilu.PARILU0, "ilu0_isai": bilu.ILU0_ISAI,
          "parilu0_isai": bilu.PARILU0_ISAI, "general_isai": bilu.GENERAL_ISAI}
    alg_bytes_per_system = 8 * nnz + 16 * n + 24
    alg_bytes = alg_bytes_per_system * B + 4 * (n + 1) + 4 * nnz

    def summary(ms):
        cvh, trh, ith = out["cv"].cpu().numpy(), out["tr"].cpu().numpy(), out["it"].cpu().numpy()
        relres = trh / bnorm
        rf, fs = out["rf"].cpu().numpy(), out["fs"].cpu().numpy()
        return {"systems_per_s": round(B / (ms * 1e-3), 1), "ms_per_step": round(ms, 4),
                "all_converged_relres_le_1e-10": bool((cvh == 1).all() and (relres <= 1e-10).all()),
                "max_relres": float(relres.max()),
                "iterations_mean": float(ith.mean()), "iterations_max": float(ith.max()),
                "refine_steps_mean": float(rf.mean()), "refine_steps_max": int(rf.max()),
                "parilu_sweeps_mean": float(fs[fs >= 0].mean()) if (fs >= 0).any() else None,
                "parilu_fallbacks": int((fs < 0).sum())}

    def run(o, F_init=None, pre=None, kernel=None, status=True):
        # the factor status is requested only from the ParILU paths (asking for
        # it routes a plain solve through the split factor + solve launches)
        par = status and o.precond in (bilu.PARILU0, bilu.PARILU0_ISAI)
        out["fs"].zero_()
        out["rf"].zero_()

        def step():
            if pre is not None:
                pre()
            bilu.solve_ex(P, dA, db, x, o, F_init, out["it"], out["cv"], out["ir"], out["tr"],
                          out["rf"], out["fs"] if par else False, stream=stream)
        return time_steps(step, args.warmup, args.steps, stream, kernel=kernel)

    def roofline(kname, kernel_ms, key, step_ms):
        # achieved = the path's algorithmic bytes per launch over the kernel's
        # own duration (HIP events on its stream, this run)
        if not kernel_ms > 0:
            return None
        ach = alg_bytes / (kernel_ms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(ach / HBM_PEAK_GBS, 5), "kernel": kname, "kernel_ms": round(kernel_ms, 4),
             "alg_bytes_per_launch": alg_bytes,
             "step_achieved": round(alg_bytes / (step_ms * 1e-3) / 1e9, 2)}
        r.update(traffic_fields(key))
        r["lds"] = lds_fields(key, kernel_ms)
        return r

    rec = {"metric": METRIC, "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "data": "synthetic (counter-based generator, pattern seed 144, value seed 20261203)"}
    base_cfg = {"batch": B, "n": n, "nnz": nnz, "levels_lower": info.levels_lower,
                "levels_upper": info.levels_upper,
                "parallelism": "one wave64 per system; persistent workgroups over the CUs"}

    if args.mode == "single":
        mixed = args.precision == "fp32_ir"
        o = bilu.opts(precond=PC[args.precond], parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE,
                      true_res_restart=1, max_restarts=3,
                      merge={"auto": bilu.MERGE_AUTO, "merged": bilu.MERGED, "segregated": bilu.SEGREGATED}[args.merge],
                      kernel={"auto": bilu.KERNEL_AUTO, "wave": bilu.KERNEL_WAVE,
                              "persistent": bilu.KERNEL_PERSISTENT}[args.kernel],
                      precision=bilu.PRECISION_FP32_IR if mixed else bilu.PRECISION_FP64,
                      inner_tol=args.inner_tol, isai_power=args.isai_power)
        ms, kms = run(o, kernel=bilu.TIMING_SOLVE)
        s = summary(ms)
        kname = ("k_mixed" if mixed else
                 "k_fast" if (args.kernel != "wave" and args.isai_power == 1 and n <= 192) else "k_solve")
        rec.update({"value": s["systems_per_s"], "ms_per_step": s["ms_per_step"],
                    "dtype": "f32+f64" if mixed else "f64",
                    "config": dict(base_cfg, workload=f"random pattern n={n} nnz={nnz}, batch {B}, "
                                                      f"{args.precond} {args.merge} {args.precision}",
                                   precond=args.precond, merge=args.merge, kernel=args.kernel,
                                   precision=args.precision, isai_power=args.isai_power),
                    "result": s,
                    "roofline": roofline(kname, kms, f"{kname} {args.precond} {args.precision} n{n} b{B}", ms)})
        rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline(pat, A, b, args.precond)
        print(json.dumps(rec), flush=True)
        return

    if args.mode == "learned":
        # the headline's learned step alone: network trained on --train-systems
        # for --train-epochs, then bilu_mlp_solve timed (both kernels' durations
        # from the library's HIP events)
        o_learn = bilu.opts(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8,
                            parilu_order=bilu.PARILU_INPLACE, precision=bilu.PRECISION_FP32_IR,
                            inner_tol=args.inner_tol, tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
        P.reserve(B)
        net = bilu.Mlp(P, hidden=512, seed=1)
        if args.train_epochs > 0 and args.train_systems > 0:
            _, At, _ = synth.batch("random", args.train_systems, first=1 << 24, n=n)
            dAt = torch.from_numpy(At).to(dev)
            dFt = torch.empty_like(dAt)
            bilu.ilu0(P, dAt, dFt)
            for ep in range(args.train_epochs):
                for s0 in range(0, dAt.shape[0], 4096):
                    net.train_step(dAt[s0:s0 + 4096], dFt[s0:s0 + 4096], 3e-4)
            del dAt, dFt

        def learned_step():
            net.solve(P, dA, db, x, o_learn, out["it"], out["cv"], out["ir"], out["tr"], out["rf"], out["fs"],
                      stream=stream)
        ms, kms = time_steps(learned_step, args.warmup, args.steps, stream, kernel=bilu.TIMING_SOLVE)
        s = summary(ms)
        _, kmf = time_steps(learned_step, 0, args.steps, stream, kernel=bilu.TIMING_FACTOR)
        rec.update({"value": s["systems_per_s"], "ms_per_step": s["ms_per_step"], "dtype": "f32+f64",
                    "config": dict(base_cfg, workload=f"random pattern n={n} nnz={nnz}, batch {B}, learned + "
                                                      f"mixed step (trained {args.train_epochs} epochs on "
                                                      f"{args.train_systems} systems)"),
                    "result": s, "kernel_ms": {"k_mixed": round(kms, 4), "k_factor_fast": round(kmf, 4)},
                    "roofline": roofline("k_mixed", kms, f"k_mixed learned n{n} b{B}", ms)})
        print(json.dumps(rec), flush=True)
        return

    # ---------------------------------------------------------------- headline
    common = dict(tol=1e-10, tol_mode=bilu.TOL_RELATIVE)
    P.reserve(B)
    copy_peak = hbm_copy_peak(dev)
    # (1) plain FP64 merged exact ILU(0) + BiCGSTAB (the persistent k_fast kernel)
    o_plain = bilu.opts(precond=bilu.ILU0, true_res_restart=1, **common)
    ms_plain, kms_plain = run(o_plain, kernel=bilu.TIMING_SOLVE)
    plain = summary(ms_plain)
    x_plain = x.clone()
    # (2) mixed precision without the network: exact ILU(0), ISAI(L, U) in FP32, FP32 inner / FP64 IR
    o_mixed = bilu.opts(precond=bilu.ILU0_ISAI, precision=bilu.PRECISION_FP32_IR,
                        inner_tol=args.inner_tol, **common)
    ms_mixed = run(o_mixed)
    mixed = summary(ms_mixed)
    x_mixed = x.clone()
    # (2a) the same in one launch: the merged kernel factors each system in LDS, then solves it
    o_mixed_m = bilu.opts(precond=bilu.ILU0_ISAI, precision=bilu.PRECISION_FP32_IR, merge=bilu.MERGED,
                          inner_tol=args.inner_tol, **common)
    mixed_merged = summary(run(o_mixed_m))
    mixed_merged["x_equals_segregated"] = bool(torch.equal(x, x_mixed))
    # (2a') FP32 general ISAI of A (FP64 generation launch, FP32 values) in the same refinement
    o_mixed_g = bilu.opts(precond=bilu.GENERAL_ISAI, precision=bilu.PRECISION_FP32_IR, inner_tol=args.inner_tol,
                          **common)
    mixed_general = summary(run(o_mixed_g))
    # (2b) the other FP64 solvers of the plain method on the persistent kernels:
    # ParILU(0) 5 sweeps merged into the solve; ILU(0) / ParILU(0) + ISAI(L, U)
    # and the general ISAI of A (generation launch, then the persistent solve)
    c1 = {}
    for name, pc, mg in (("parilu0_5_sweeps", bilu.PARILU0, bilu.MERGE_AUTO),
                         ("parilu0_5_sweeps_merged", bilu.PARILU0, bilu.MERGED),
                         ("ilu0_isai", bilu.ILU0_ISAI, bilu.MERGE_AUTO),
                         ("parilu0_5_sweeps_isai", bilu.PARILU0_ISAI, bilu.MERGE_AUTO),
                         ("general_isai", bilu.GENERAL_ISAI, bilu.MERGE_AUTO)):
        o_v = bilu.opts(precond=pc, parilu_sweeps=5, true_res_restart=1, kernel=bilu.KERNEL_PERSISTENT, merge=mg,
                        **common)
        ms_v, kms_v = run(o_v, kernel=bilu.TIMING_SOLVE, status=False)
        c1[name] = dict(summary(ms_v), solve_kernel_ms=round(kms_v, 4))
    for name, pc in (("ilu0_segregated", bilu.ILU0), ("parilu0_5_sweeps_segregated", bilu.PARILU0)):
        o_seg = bilu.opts(precond=pc, parilu_sweeps=5, true_res_restart=1, merge=bilu.SEGREGATED, **common)
        ms_v, kms_v = run(o_seg, kernel=bilu.TIMING_SOLVE, status=False)
        c1[name] = dict(summary(ms_v), solve_kernel_ms=round(kms_v, 4))
    # (3) learned + mixed: train the warm start (untimed), then time inference + solve
    o_learn = bilu.opts(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8,
                        parilu_order=bilu.PARILU_INPLACE, precision=bilu.PRECISION_FP32_IR,
                        inner_tol=args.inner_tol, **common)
    # held-out systems for the training curve (ParILU sweeps to 1e-8 from the network's guess)
    _, Ah, bh = synth.batch("random", 8192, first=(1 << 26), n=n)
    dAh, dbh = torch.from_numpy(Ah).to(dev), torch.from_numpy(bh).to(dev)
    xh = torch.empty_like(dbh)

    def held_sweeps(net=None):
        fs = torch.empty(dAh.shape[0], dtype=torch.int32, device=dev)
        if net is None:
            bilu.solve_ex(P, dAh, dbh, xh, o_learn, factor_status=fs)
        else:
            net.solve(P, dAh, dbh, xh, o_learn, factor_status=fs)
        torch.cuda.synchronize()
        return round(float(fs.float().mean().item()), 4)

    t0 = time.time()
    net = bilu.Mlp(P, hidden=512, seed=1)
    net0 = bilu.Mlp(P, hidden=512, seed=1)   # untrained: the base guess
    chunk = 65536
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    tr_A, tr_F = [], []
    for c0 in range(0, args.train_systems, chunk):
        _, At, _ = synth.batch("random", min(chunk, args.train_systems - c0), first=(1 << 24) + c0, n=n)
        dAt = torch.from_numpy(At).to(dev)
        dFt = torch.empty_like(dAt)
        bilu.ilu0(P, dAt, dFt)           # exact FP64 ILU(0) labels from the library itself
        tr_A.append(dAt)
        tr_F.append(dFt)
    torch.cuda.synchronize()
    gen_train_s = time.time() - t0
    held_from_A, held_base = held_sweeps(None), held_sweeps(net0)
    t1 = time.time()
    mb, steps_done, curve = 4096, 0, []
    total_steps = max(1, args.train_epochs * args.train_systems // mb)
    every = max(1, args.train_epochs // 6)
    for ep in range(args.train_epochs):
        for ci in range(len(tr_A)):
            for s0 in range(0, tr_A[ci].shape[0], mb):
                lr = 3e-4 * 0.5 * (1.0 + math.cos(math.pi * steps_done / total_steps))   # cosine decay
                net.train_step(tr_A[ci][s0:s0 + mb], tr_F[ci][s0:s0 + mb], lr, loss)
                steps_done += 1
        if ep % every == 0 or ep == args.train_epochs - 1:
            curve.append({"epoch": ep + 1, "loss": round(loss.item(), 6), "held_out_sweeps": held_sweeps(net)})
    torch.cuda.synchronize()
    train_s = time.time() - t1
    del tr_A, tr_F
    # how accurate a warm start must be to save sweeps: the exact ILU(0) factors of
    # the held-out systems with relative noise eps on every entry
    dFh = torch.empty_like(dAh)
    bilu.ilu0(P, dAh, dFh)
    gen = torch.Generator(device=dev).manual_seed(1234)
    noise = torch.randn(dFh.shape, generator=gen, device=dev, dtype=torch.float64)
    fsh = torch.empty(dAh.shape[0], dtype=torch.int32, device=dev)
    accuracy = {}
    for eps in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
        bilu.solve_ex(P, dAh, dbh, xh, o_learn, F_init=dFh * (1.0 + eps * noise), factor_status=fsh)
        torch.cuda.synchronize()
        accuracy[f"{eps:g}"] = round(float(fsh.float().mean().item()), 4)
    del dFh, noise
    guess = torch.empty_like(dA)
    # (3a) split: bilu_mlp_infer writes the FP64 guess, bilu_solve_ex reads it
    ms_split = run(o_learn, F_init=guess, pre=lambda: net.infer(dA, guess, stream=stream))
    split = summary(ms_split)
    x_split = x.clone()
    # (3) fused (bilu_mlp_solve): the factor kernel forms the guess from the
    # network's raw outputs; the same results bit for bit.  The solve kernel
    # (k_mixed, the step's longest) is timed with the library's HIP events.
    out["fs"].zero_()
    out["rf"].zero_()

    def fused_step():
        net.solve(P, dA, db, x, o_learn, out["it"], out["cv"], out["ir"], out["tr"], out["rf"], out["fs"],
                  stream=stream)
    ms_learn, kms_learn = time_steps(fused_step, args.warmup, args.steps, stream, kernel=bilu.TIMING_SOLVE)
    learned = summary(ms_learn)
    fused_equal = bool(torch.equal(x, x_split))
    x_learn = x.clone()
    # (3m) the learned step with the merged kernel (warm start, sweeps, ISAI and solve in one launch)
    o_learn_m = bilu.opts(precond=bilu.PARILU0_ISAI, parilu_sweeps=40, parilu_stop_tol=1e-8,
                          parilu_order=bilu.PARILU_INPLACE, precision=bilu.PRECISION_FP32_IR, merge=bilu.MERGED,
                          inner_tol=args.inner_tol, **common)
    out["fs"].zero_()
    out["rf"].zero_()

    def merged_step():
        net.solve(P, dA, db, x, o_learn_m, out["it"], out["cv"], out["ir"], out["tr"], out["rf"], out["fs"],
                  stream=stream)
    learned_merged = summary(time_steps(merged_step, args.warmup, args.steps, stream))
    learned_merged["x_equals_segregated"] = bool(torch.equal(x, x_learn))
    _, kms_factor = time_steps(fused_step, 0, max(2, args.steps // 2), stream, kernel=bilu.TIMING_FACTOR)
    # (3b) the same ParILU path from the standard start (no network), for the gain
    ms_nolearn = run(o_learn)
    nolearn = summary(ms_nolearn)
    # (3c) the untrained network (zero output layer: the base guess a_ij, a_ij / a_jj)
    out["fs"].zero_()
    out["rf"].zero_()

    def untrained_step():
        net0.solve(P, dA, db, x, o_learn, out["it"], out["cv"], out["ir"], out["tr"], out["rf"], out["fs"],
                   stream=stream)
    untrained = summary(time_steps(untrained_step, args.warmup, args.steps, stream))
    net0.close()
    ms_infer = time_steps(lambda: net.infer(dA, guess, stream=stream), 1, args.steps, stream)

    # FP32-IR solution error against dense FP64 solves, on a sample of systems
    smp = np.arange(0, B, max(1, B // 256))[:256]
    xs = np.stack([np.linalg.solve(pat.dense(A[i]), b[i]) for i in smp])
    def err(xg):
        xg = xg.cpu().numpy()[smp]
        e = np.linalg.norm(xg - xs, axis=1) / np.linalg.norm(xs, axis=1)
        return {"max": float(e.max()), "mean": float(e.mean())}
    c2 = {"sample": f"{len(smp)} systems (every {B // len(smp)}th), reference numpy.linalg.solve in FP64",
          "plain_fp64": err(x_plain), "mixed_ilu0_isai": err(x_mixed), "learned_mixed": err(x_learn),
          "refine_steps_mean": {"mixed_ilu0_isai": mixed["refine_steps_mean"],
                                "learned_mixed": learned["refine_steps_mean"]},
          "refine_steps_max": {"mixed_ilu0_isai": mixed["refine_steps_max"],
                               "learned_mixed": learned["refine_steps_max"]}}

    dom = "k_mixed" if kms_learn >= kms_factor else "k_factor_fast"
    rec.update({
        "value": learned["systems_per_s"], "ms_per_step": learned["ms_per_step"], "dtype": "f32+f64",
        "config": dict(base_cfg,
                       workload=(f"random pattern n={n} nnz={nnz}, batch {B}: MLP warm start (BF16) -> "
                                 f"in-place ParILU(0) to factor residual 1e-8 (exact ILU(0) fallback) -> "
                                 f"ISAI(L,U) FP32 -> FP32 BiCGSTAB (inner tol {args.inner_tol:g}, adaptive) in "
                                 f"FP64 refinement to relres 1e-10"),
                       variant="learned+mixed"),
        "speedup_vs_plain_fp64": round(ms_plain / ms_learn, 3),
        "variants": {"plain_fp64_ilu0_merged": plain, "mixed_ilu0_isai": mixed,
                     "mixed_ilu0_isai_merged_kernel": mixed_merged, "mixed_general_isai": mixed_general,
                     "learned_mixed_parilu0_isai": learned, "learned_mixed_parilu0_isai_merged_kernel": learned_merged,
                     "learned_mixed_parilu0_isai_split": split,
                     "mixed_parilu0_isai_no_warm_start": nolearn,
                     "untrained_network_base_guess_mixed_parilu0_isai": untrained},
        "fp64_solver_variants": c1,
        "fused_x_equals_split_x": fused_equal,
        "mlp_inference_ms": round(ms_infer, 4),
        "kernel_ms": {"k_mixed": round(kms_learn, 4), "k_factor_fast": round(kms_factor, 4),
                      "k_fast_plain_fp64": round(kms_plain, 4),
                      "source": "HIP events recorded by the library around each launch on its stream"},
        "training": {"systems": args.train_systems, "epochs": args.train_epochs, "minibatch": mb,
                     "schedule": "Adam, lr 3e-4 cosine to 0", "label_generation_s": round(gen_train_s, 2),
                     "train_s": round(train_s, 2),
                     "held_out": {"systems": int(dAh.shape[0]), "sweeps_from_A": held_from_A,
                                  "sweeps_untrained_base_guess": held_base, "curve": curve,
                                  "sweeps_from_exact_factors_with_relative_noise": accuracy}},
        "all_converged_relres_le_1e-10": learned["all_converged_relres_le_1e-10"],
        "roofline": roofline(dom, kms_learn if dom == "k_mixed" else kms_factor,
                             f"{dom} learned n{n} b{B}", ms_learn),
        "roofline_plain_fp64": roofline("k_fast", kms_plain, f"k_fast ilu0 fp64 n{n} b{B}", ms_plain),
        "hbm_copy_peak_gbs": copy_peak,
        "fp32_ir_error_vs_fp64": c2,
        "gen_seconds": round(gen_s, 2),
    })
    rec["uneven"] = None if args.no_uneven else uneven_bench(args, dev)
    if not args.no_uneven:
        other = "ilu0" if args.uneven_precond == "ilu0_isai" else "ilu0_isai"
        u2 = uneven_bench(args, dev, precond=other)
        rec["uneven"]["other_precond"] = {k: u2[k] for k in ("value", "ms_per_step", "all_converged_relres_le_1e-10",
                                                             "tuned", "config")}
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline(pat, A, b, "ilu0")
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
