#!/usr/bin/env python3
"""Benchmark: FP64 SpMV effective HBM bandwidth on one MI355X.

Workload (one step): y = A*x with the library's adaptive CSR kernel on
  (a) the 27-point stencil on a 64^3 grid (n=262,144, nnz=6,859,000) and
  (b) the power-law matrix (n=2^20, lognormal row lengths mean 20, sigma 1.2,
      cap 5000, ~21M nnz),
synthetic data from the library's deterministic generators, resident in HBM.
Each matrix is replicated so that the copies together exceed 512 MB and every
launch reads a copy that is not in the 256 MB Infinity Cache (cold HBM).

value = sum of CSR-equivalent algorithmic bytes (12*nnz + 4*(n+1) + 8*ncols
+ 8*n per SpMV) over the timed steps / timed wall time, in GB/s.

Prints one JSON line. `--gpus N` with N > 1 is refused: this build measures
one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20
METRIC = "SpMV fp64 eff. GB/s (% HBM peak); SpMM fp16 fwd+bwd GFLOP/s vs plain fp32"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--method", default="adaptive")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the CPU-baseline sample")
    ap.add_argument("--warm", action="store_true", help="single copy (cache-warm) timing")
    return ap.parse_args()


def _copies(A, warm: bool) -> int:
    if warm:
        return 1
    return max(2, math.ceil(2 * MALL_BYTES / A.csr_bytes()))


def cpu_baseline(mats, xs, seconds: float):
    """The oracle's restatement of the reference CPU loop (src/ref.cpp:26-37),
    row-partitioned over the host's threads (bit-identical per row), on a
    bounded sample: repeated SpMVs over both matrices for ~`seconds`."""
    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    O.lib()
    nbytes, t_total, reps = 0, 0.0, 0
    while t_total < seconds or reps < 1:
        for A, x in zip(mats, xs):
            t0 = time.perf_counter()
            O.spmv_f64(A.rowptr, A.col, A.val, x, threads=threads)
            t_total += time.perf_counter() - t0
            nbytes += A.csr_bytes()
        reps += 1
    return {"value": nbytes / t_total / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{reps} x (SpMV 27pt-64^3 + SpMV powerlaw-1M), {t_total:.1f} s, "
                      f"{threads} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU sharding plan is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from benchsp_amd import host
    from benchsp_amd.device import CsrSpmv, DeviceCsr

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    mats = [("27pt-64^3", host.stencil27_64()), ("powerlaw-1M", host.powerlaw_1m())]
    xs = [host.uniform(A.ncols, seed=100 + i) for i, (_, A) in enumerate(mats)]

    work = []  # per matrix: list of (op), x, y, bytes
    for (name, A), x in zip(mats, xs):
        ncopy = _copies(A, args.warm)
        ops = [CsrSpmv(DeviceCsr.from_host(A, dev), args.method) for _ in range(ncopy)]
        xd = torch.from_numpy(x).to(dev)
        yd = torch.empty(A.nrows, dtype=torch.float64, device=dev)
        work.append(dict(name=name, A=A, ops=ops, x=xd, y=yd, bytes=A.csr_bytes(), ncopy=ncopy))
    torch.cuda.synchronize()

    def step(i, evs=None):
        for wi, w in enumerate(work):
            op = w["ops"][i % w["ncopy"]]
            if evs is not None:
                evs[wi][0].record()
            op(w["x"], w["y"])
            if evs is not None:
                evs[wi][1].record()

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()

    # timed region: wall clock over K steps (bracketed by synchronize)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # per-launch kernel durations with HIP events on the launch stream
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in work]
          for _ in range(args.steps)]
    for i in range(args.steps):
        step(args.warmup + i, ev[i])
    torch.cuda.synchronize()
    kern_ms = [[ev[i][wi][0].elapsed_time(ev[i][wi][1]) for i in range(args.steps)]
               for wi in range(len(work))]

    step_bytes = sum(w["bytes"] for w in work)
    value = step_bytes * args.steps / wall / 1e9
    per = {}
    tot_b, tot_t = 0.0, 0.0
    for wi, w in enumerate(work):
        ms = float(np.mean(kern_ms[wi]))
        tot_b += w["bytes"]
        tot_t += ms
        per[w["name"]] = {"nrows": w["A"].nrows, "nnz": w["A"].nnz, "bytes": w["bytes"],
                          "kernel_us": round(ms * 1e3, 2),
                          "eff_GBs": round(w["bytes"] / (ms * 1e-3) / 1e9, 1),
                          "pct_hbm_peak": round(100 * w["bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1),
                          "copies": w["ncopy"]}
    achieved = tot_b / (tot_t * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FP64 CSR SpMV: 27-pt stencil 64^3 + power-law n=2^20 (lognormal "
                               "mean 20, sigma 1.2, cap 5000)",
                   "method": args.method, "cache": "warm" if args.warm else "cold (rotated copies)",
                   "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 1),
                   "matrices": per},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "csr_adaptive_f64" if args.method == "adaptive" else args.method},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline([A for _, A in mats], xs, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
