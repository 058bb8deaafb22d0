"""Benchmark: systems solved per second to relres <= 1e-10 on a batch of
65536 sparse systems (n=144, one shared random pattern, ~2000 nnz), one GPU.

One step = one batched solve (preconditioner setup + BiCGSTAB + true-residual
epilogue) over the whole batch, inputs resident in HBM.  Prints one JSON
line (see the repository's bench contract).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "merging-ilu0_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--n", type=int, default=144)
    ap.add_argument("--precond", default="ilu0",
                    choices=["ilu0", "parilu0", "ilu0_isai", "parilu0_isai", "general_isai"])
    ap.add_argument("--merge", default="merged", choices=["merged", "segregated"])
    ap.add_argument("--kernel", default="auto", choices=["auto", "wave", "persistent"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir"],
                    help="fp32_ir: FP32 preconditioner and inner BiCGSTAB inside FP64 refinement")
    ap.add_argument("--inner-tol", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a 1 GiB device copy (HBM counter calibration)")
    return ap.parse_args()


def measured_traffic(key):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json), each
    scaled by what the counter reads for a 1 GiB device copy; None when that
    workload has not been profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    return None if rec is None else rec["traffic_bytes_per_launch"]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pat, A, b, precond):
    """The oracle (CPU restatement of the reference path, same algorithm and
    stopping rule) timed on this host: a bounded prefix of the same batch on
    all threads, plus a single-core figure on a smaller prefix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    code = {"ilu0": O.ILU0, "parilu0": O.PARILU0, "ilu0_isai": O.ILU0_ISAI,
            "parilu0_isai": O.PARILU0_ISAI, "general_isai": O.GENERAL_ISAI}[precond]
    oo = O.opts(tol=1e-10, tol_mode=O.TOL_RELATIVE, true_res_restart=1)
    threads = max(1, min(os.cpu_count() or 1, 16))
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
            "sample": f"{s_all} systems (the whole batch), {threads} threads, {dt_all:.2f} s wall",
            "single_core_value": round(s_one / dt_one, 1),
            "single_core_sample": f"first {s_one} systems, 1 thread, {dt_one:.2f} s",
            "cpu_model": cpu_model()}


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

    t0 = time.time()
    pat, A, b = synth.batch("random", args.batch, n=args.n)
    gen_s = time.time() - t0
    P = bilu.Pattern.from_synth(pat)
    info = P.info()
    dA = torch.from_numpy(A).to(dev)
    db = torch.from_numpy(b).to(dev)
    x = torch.empty_like(db)
    B, n, nnz = args.batch, pat.n, pat.nnz
    it = torch.empty(B, dtype=torch.float32, device=dev)
    cv = torch.empty(B, dtype=torch.int32, device=dev)
    ir = torch.empty(B, dtype=torch.float64, device=dev)
    tr = torch.empty(B, dtype=torch.float64, device=dev)
    pc = {"ilu0": bilu.ILU0, "parilu0": bilu.PARILU0, "ilu0_isai": bilu.ILU0_ISAI,
          "parilu0_isai": bilu.PARILU0_ISAI, "general_isai": bilu.GENERAL_ISAI}[args.precond]
    merge = bilu.MERGED if args.merge == "merged" else bilu.SEGREGATED
    kern = {"auto": bilu.KERNEL_AUTO, "wave": bilu.KERNEL_WAVE,
            "persistent": bilu.KERNEL_PERSISTENT}[args.kernel]
    mixed = args.precision == "fp32_ir"
    o = bilu.opts(precond=pc, parilu_sweeps=5, tol=1e-10, tol_mode=bilu.TOL_RELATIVE,
                  true_res_restart=1, max_restarts=3, merge=merge, kernel=kern,
                  precision=bilu.PRECISION_FP32_IR if mixed else bilu.PRECISION_FP64,
                  inner_tol=args.inner_tol)
    stream = torch.cuda.current_stream()
    refine = torch.zeros(B, dtype=torch.int32, device=dev)
    fstat = torch.zeros(B, dtype=torch.int32, device=dev)

    def step():
        if mixed:
            bilu.solve_ex(P, dA, db, x, o, None, it, cv, ir, tr, refine, fstat, stream=stream)
        else:
            bilu.solve(P, dA, db, x, o, it, cv, ir, tr, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / max(args.steps, 1)
    value = B / (ms_per_step * 1e-3)

    # acceptance: every system converged with true relres <= 1e-10
    cvh, trh, ith = cv.cpu().numpy(), tr.cpu().numpy(), it.cpu().numpy()
    relres = trh / np.linalg.norm(b, axis=1)
    all_ok = bool((cvh == 1).all() and (relres <= 1e-10).all())

    # roofline of the dominant kernel (the merged solve is one launch per step)
    alg_bytes_per_system = 8 * nnz + 16 * n + 24
    alg_bytes = alg_bytes_per_system * B + 4 * (n + 1) + 4 * nnz
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    kname = ("k_mixed" if mixed else
             "k_fast" if (args.merge == "merged" and args.kernel != "wave") else None)
    traffic = measured_traffic(f"{kname} {args.precond} {args.merge} n{n} b{B}") if kname else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": ("k_factor_fast + k_mixed (FP32 inner, FP64 refinement)" if mixed else
                           ("k_solve (one wave per system)" if args.kernel == "wave" else
                            "k_fast (persistent multi-wave)") if args.merge == "merged" else "k_factor + solve"),
                "alg_bytes_per_launch": alg_bytes}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pat, A, b, args.precond)

    rec = {
        "metric": "systems solved/s at relres<=1e-10, n=144 batch 65536: learned+mixed vs plain FP64",
        "value": round(value, 1),
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (counter-based generator, pattern seed 144, value seed 20261203)",
        "config": {"workload": (f"random pattern n={n} nnz={nnz}, batch {B}, {args.precond} factor "
                                f"(FP64, rounded to FP32) + FP32 BiCGSTAB (inner tol {args.inner_tol:g}) "
                                f"inside FP64 iterative refinement, relres<=1e-10" if mixed else
                                f"random pattern n={n} nnz={nnz}, batch {B}, exact ILU(0) merged "
                                f"factor+BiCGSTAB, FP64, relres<=1e-10 (variant: plain FP64 "
                                f"{args.precond} {args.merge})"),
                   "batch": B, "n": n, "nnz": nnz, "precond": args.precond, "merge": args.merge,
                   "levels_lower": info.levels_lower, "levels_upper": info.levels_upper,
                   "kernel": args.kernel, "precision": args.precision,
                   "inner_tol": args.inner_tol if mixed else None,
                   "parallelism": "one wave64 per system; batch sharded over CUs"},
        "all_converged_relres_le_1e-10": all_ok,
        "max_relres": float(relres.max()),
        "iterations": {"min": float(ith.min()), "mean": float(ith.mean()), "max": float(ith.max())},
        "refine_steps": ({"mean": float(refine.float().mean()), "max": int(refine.max())}
                         if mixed else None),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gen_seconds": round(gen_s, 2),
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
