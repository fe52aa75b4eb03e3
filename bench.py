#!/usr/bin/env python3
"""Benchmark of the batched Epanechnikov KDE hot path on one MI355X.

Workload (default): the all-pairs mutual-information matrix of 64 columns x
200,000 frames (seeded 64x64 N(0,1) mixing of independent 2-component
Gaussian mixtures, z-scored per column, float32), bandwidth 0.15 per dim:
64 1-D + 2016 2-D densities in one batched persistent launch, MI reduced on
the device in FP64.  A step is one full MI matrix from device-resident
coordinates (binning, sort, density launch, reductions, 64x64 result to host).

Printed JSON line: value = executed pair evaluations per second (Gpair-evals/s)
over the whole step; the dense-equivalent rate, the plain per-pair GPU path
(the reference's 4096-call driver order, extrapolated from timed calls) and
the host OpenMP baseline (oracle restatement of densities_omp.cpp on a timed
subset, extrapolated) are reported beside it.  Roofline: FP32 VALU, achieved
= algorithmic FLOPs (4d per executed pair) / duration of the persistent
density kernel measured with HIP events.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_FP32_TFLOPS = 157.3
FLOPS_PER_PAIR = (4, 8, 12)  # d = 1, 2, 3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def time_call(fn, reps: int = 3) -> float:
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline_mi(X: np.ndarray, h: float, threads: int) -> dict:
    """Host OpenMP path (the oracle's restatement of densities_omp.cpp: same
    sort, O(N^2) linear window scans, FP32 window sums, same schedules) timed
    on one 1-D and one 2-D density of the workload, extrapolated to the 4096
    calls of mi.cpp's driver loop (2080 1-D + 2016 2-D)."""
    import oracle

    oracle.set_threads(threads)
    n = X.shape[1]
    t1 = time_call(lambda: oracle.density_omp_packed(X[0:1], n, [h]), reps=1)
    log(f"cpu 1-D density {t1:.2f} s")
    t2 = time_call(lambda: oracle.density_omp_packed(X[0:2], n, [h, h]), reps=1)
    log(f"cpu 2-D density {t2:.2f} s")
    total = 2080 * t1 + 2016 * t2
    return {
        "value": total,
        "unit": "s per MI matrix (extrapolated)",
        "cores": threads,
        "kind": "port",
        "sample": f"one 1-D and one 2-D density at N={n} ({t1:.2f} s, {t2:.2f} s), scaled by mi.cpp's 2080+2016 calls",
    }


def run_mi(args) -> dict:
    from entropical_amd import Context, DeviceArray, combined_densities, datasets, mi_matrix

    n_cols, n_rows, h = 64, 200_000, 0.15
    log("generating 64 x 200000 mixed columns")
    X = datasets.mixed_columns(n_cols, n_rows)
    hs = [h] * n_cols
    ctx = Context(0)
    dx = DeviceArray(ctx, X)
    log("warmup")
    for _ in range(args.warmup):
        mi_matrix(ctx, dx, hs)
    ctx.sync()
    log(f"timing {args.steps} steps")
    dens_ms, pairs_by_d = [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mi = mi_matrix(ctx, dx, hs)
        st = ctx.last_stats()
        dens_ms.append(st["density_ms"])
        pairs_by_d = st["pairs_by_d"]
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    pairs = sum(pairs_by_d)
    flops = sum(p * f for p, f in zip(pairs_by_d, FLOPS_PER_PAIR))
    kern_ms = float(np.mean(dens_ms))
    achieved = flops / (kern_ms * 1e-3) / 1e12
    dense_pairs = 2080 * n_rows * n_rows
    log(f"MI matrix {ms_step:.1f} ms/step, density kernel {kern_ms:.1f} ms, {pairs:.3e} pairs")
    assert np.all(np.isfinite(mi)) and np.allclose(mi, mi.T)

    res = {
        "metric": "Gpair-evals/s (batched 64-column MI matrix, 200k frames)",
        "value": pairs / (elapsed / args.steps) / 1e9,
        "unit": "Gpair-evals/s",
        "ms_per_step": ms_step,
        "dtype": "f32",
        "config": {
            "workload": "mi_matrix_64col_200000frames_h0.15",
            "jobs": "64 1-D + 2016 2-D densities, one persistent launch",
            "n_cols": n_cols,
            "n_rows": n_rows,
            "bandwidth": h,
            "method": "cell list",
        },
        "pairs_executed": pairs,
        "pairs_by_dim": pairs_by_d,
        "dense_equiv_gpair_per_s": dense_pairs / (elapsed / args.steps) / 1e9,
        "roofline": {
            "bound": "valu",
            "kernel": "k_density (persistent)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "kernel_ms": kern_ms,
            "flops_per_pair": "4d (d=1: 4, d=2: 8)",
        },
    }
    if not args.no_plain:
        # plain per-pair GPU path: the reference's driver order, each call
        # with its H2D, density launch and D2H (4096 calls for 64 columns)
        log("timing plain per-pair path (one call each, extrapolated)")
        plain = {}
        for method in ("brute", "cell"):
            combined_densities(X, n_rows, [0], [h], method=method, ctx=ctx)
            t1 = time_call(lambda: combined_densities(X, n_rows, [0], [h], method=method, ctx=ctx))
            t2 = time_call(lambda: combined_densities(X, n_rows, [0, 1], [h, h], method=method, ctx=ctx))
            plain[method] = {"t_1d_ms": 1e3 * t1, "t_2d_ms": 1e3 * t2, "mi_matrix_ms": 1e3 * (2080 * t1 + 2016 * t2)}
        res["plain_path"] = plain
        res["speedup_vs_plain_brute"] = plain["brute"]["mi_matrix_ms"] / ms_step
        res["speedup_vs_plain_cell"] = plain["cell"]["mi_matrix_ms"] / ms_step
    dx.free()
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cb = cpu_baseline_mi(X, h, args.cpu_threads or host_cores())
        res["cpu_baseline"] = cb
        res["speedup_vs_cpu"] = cb["value"] * 1e3 / ms_step
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py measures one GPU only: this build has no multi-GPU path "
            "(jobs would be sharded by cost across GPUs with one input broadcast "
            "and one gather of per-job scalars); run with --gpus 1.",
            file=sys.stderr,
        )
        return 2
    res = run_mi(args)
    line = {
        "metric": res.pop("metric"),
        "value": res.pop("value"),
        "unit": res.pop("unit"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": res.pop("dtype"),
        "data": "synthetic (seeded generator)",
        "config": res.pop("config"),
    }
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
