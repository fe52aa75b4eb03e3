#!/usr/bin/env python3
"""Benchmark of the batched Epanechnikov KDE hot path on one MI355X.

Default workload: the all-pairs mutual-information matrix of 64 columns x
200,000 frames (seeded 64x64 N(0,1) mixing of independent 2-component
Gaussian mixtures, z-scored per column, float32), bandwidth 0.15 per dim:
64 1-D + 2016 2-D densities in one batched persistent launch, MI reduced on
the device in FP64.  A step is one full MI matrix from device-resident
coordinates (planning, binning, sort, tiles, density launch, reductions,
64x64 result to host).

Other workloads (--workload): te (32 coupled AR(1) series x 100,000 frames,
h=0.2, lags 1,2,5,10), uneven (500 trajectories of 2,000-500,000 frames,
d 1-3), entropy1d (N=100,000 three-component mixture, h=0.1).

Printed JSON line: value = candidate pair evaluations per second
(Gpair-evals/s) over the whole step: the (query, candidate) pairs of every
tile's cell-list window, evaluated one by one or, for the boxes of a 1-D job
and the cells of a 2-D job that lie inside the support of every query of a
tile, summed exactly from their moments (both counts are reported); the
dense-equivalent rate, the plain
per-pair GPU path (the reference's per-call driver order, extrapolated from
timed calls) and the host OpenMP baseline (oracle restatement of
densities_omp.cpp on timed calls, extrapolated) are reported beside it.
Roofline: FP32 VALU; achieved = algorithmic FLOPs (4d per pair evaluated one
by one, 6 per (query, box) and 25 per (query, cell) moment term) /
duration of the persistent density kernel, measured with HIP events on the
engine's stream.  The tile shape is chosen by the engine's autotuner during
warmup (untimed).
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
# (query, box) term of a 1-D box inside every query's support, which stands
# for all of the box's pairs: d = x - r, t = B - A d, c + t d, accumulate; and
# (query, cell) term of a 2-D cell: 2 offsets, 2 x (scale, 1 - k d^2), 3
# inner FMA pairs, 1 outer FMA pair, accumulate
MOMENT_TERM_FLOPS = (6, 25)


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


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(workload: str):
    """HBM bytes per k_density launch for this workload from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by the factor the same counter showed on a copy of known size),
    or None when no summary exists for the workload."""
    path = os.path.join(PROFILES, f"k_density_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                kv[k.strip()] = v.strip()
    try:
        return int(float(kv["traffic_bytes"]))
    except (KeyError, ValueError):
        return None


def evaluated_pairs(st) -> list:
    """Pairs evaluated one by one per job dimension: the candidate pairs less
    those summed from 1-D box and 2-D cell moments."""
    by_d = list(st["pairs_by_d"])
    for d, m in enumerate(st["moment_pairs_by_d"]):
        by_d[d] -= m
    return by_d


def roofline(st, kern_ms: float, workload: str) -> dict:
    """Algorithmic FLOP rate of k_density: 4d FLOPs per pair evaluated one by
    one plus MOMENT_TERM_FLOPS per (query, box) or (query, cell) moment term."""
    flops = sum(p * f for p, f in zip(evaluated_pairs(st), FLOPS_PER_PAIR)) + sum(
        t * f for t, f in zip(st["moment_terms_by_d"], MOMENT_TERM_FLOPS))
    achieved = flops / (kern_ms * 1e-3) / 1e12
    # the rate at which one-by-one evaluation of every candidate pair would
    # have had to run to finish in the same time (not a hardware rate)
    pair_equiv = sum(p * f for p, f in zip(st["pairs_by_d"], FLOPS_PER_PAIR)) / (kern_ms * 1e-3) / 1e12
    return {
        "bound": "valu",
        "kernel": "k_density (persistent)",
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "traffic": measured_traffic(workload),
        "kernel_ms": kern_ms,
        "pair_equivalent_tflops": pair_equiv,
        "flops_per_pair": "4d per pair evaluated one by one (d=1: 4, d=2: 8, d=3: 12); "
                          "6 per (query, box) term of 1-D box moments, 25 per (query, cell) term of 2-D "
                          "cell moments",
    }


def copy_calibration() -> dict:
    """A plain streaming kernel of known traffic (1 GiB read + 1 GiB written
    per launch, two launches) to scale the FETCH_SIZE / WRITE_SIZE counters
    of a rocprofv3 --pmc run of this same command."""
    import torch

    n = 256 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 2
    return {"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 2,
            "ms_per_launch": 1e3 * dt}


def timed_steps(ctx, step, steps: int, warmup: int):
    for _ in range(warmup):
        step()
    ctx.sync()
    dens_ms, util = [], []
    st = None
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        st = ctx.last_stats()
        dens_ms.append(st["density_ms"])
        util.append(sum(evaluated_pairs(st)) / max(1, st["lane_pairs"]))
    ctx.sync()
    elapsed = time.perf_counter() - t0
    return elapsed / steps, float(np.mean(dens_ms)), st, float(np.mean(util))


def cpu_omp_time(packed: np.ndarray, n: int, h, threads: int) -> float:
    import oracle

    oracle.set_threads(threads)
    return time_call(lambda: oracle.density_omp_packed(packed, n, h), reps=1)


# ------------------------------------------------------------------ MI
def run_mi(args) -> dict:
    from entropical_amd import Context, DeviceArray, combined_densities, datasets, mi_autotune, mi_matrix

    n_cols, n_rows, h = 64, 200_000, 0.15
    log("generating 64 x 200000 mixed columns")
    X = datasets.mixed_columns(n_cols, n_rows)
    hs = [h] * n_cols
    ctx = Context(0)
    dx = DeviceArray(ctx, X)
    tun = mi_autotune(ctx, dx, hs) if not args.no_tune else ctx.get_tuning()
    log(f"tuning (q, m0, mr, group_h) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: mi_matrix(ctx, dx, hs), args.steps, args.warmup)
    mi = mi_matrix(ctx, dx, hs)
    assert np.all(np.isfinite(mi)) and np.allclose(mi, mi.T)
    pairs = sum(st["pairs_by_d"])
    log(f"MI matrix {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {pairs:.3e} pairs")
    res = {
        "metric": "Gpair-evals/s (batched 64-column MI matrix, 200k frames)",
        "value": pairs / sec / 1e9,
        "unit": "Gpair-evals/s",
        "ms_per_step": 1e3 * sec,
        "config": {
            "workload": "mi_matrix_64col_200000frames_h0.15",
            "jobs": "64 1-D + 2016 2-D densities, one persistent launch",
            "n_cols": n_cols, "n_rows": n_rows, "bandwidth": h, "method": "cell list",
            "tuning_q_m0_mr_group": list(tun),
        },
        "pairs_candidate": pairs, "pairs_by_moments": st["moment_pairs"], "moment_terms": st["moment_terms"],
        "pairs_by_dim": st["pairs_by_d"],
        "lane_utilisation": util,
        "dense_equiv_gpair_per_s": 2080 * n_rows * n_rows / sec / 1e9,
        "roofline": roofline(st, kern_ms, "mi"),
    }
    if not args.no_plain:
        log("timing plain per-pair path (one call each, extrapolated to mi.cpp's 4096 calls)")
        plain = {}
        for method in ("brute", "cell"):
            combined_densities(X, n_rows, [0], [h], method=method, ctx=ctx)
            t1 = time_call(lambda: combined_densities(X, n_rows, [0], [h], method=method, ctx=ctx))
            t2 = time_call(lambda: combined_densities(X, n_rows, [0, 1], [h, h], method=method, ctx=ctx))
            plain[method] = {"t_1d_ms": 1e3 * t1, "t_2d_ms": 1e3 * t2, "mi_matrix_ms": 1e3 * (2080 * t1 + 2016 * t2)}
        res["plain_path"] = plain
        res["speedup_vs_plain_brute"] = plain["brute"]["mi_matrix_ms"] / (1e3 * sec)
        res["speedup_vs_plain_cell"] = plain["cell"]["mi_matrix_ms"] / (1e3 * sec)
    dx.free()
    ctx.close()
    if not args.no_cpu_baseline:
        th = args.cpu_threads or host_cores()
        t1 = cpu_omp_time(X[0:1], n_rows, [h], th)
        log(f"cpu 1-D density {t1:.2f} s")
        t2 = cpu_omp_time(X[0:2], n_rows, [h, h], th)
        log(f"cpu 2-D density {t2:.2f} s")
        total = 2080 * t1 + 2016 * t2
        res["cpu_baseline"] = {
            "value": total, "unit": "s per MI matrix (extrapolated)", "cores": th, "kind": "port",
            "sample": f"one 1-D and one 2-D density at N={n_rows} ({t1:.2f} s, {t2:.2f} s) with the oracle's "
                      "restatement of densities_omp.cpp, scaled by mi.cpp's 2080 + 2016 calls",
        }
        res["speedup_vs_cpu"] = total / sec
    return res


# ------------------------------------------------------------------ TE
def run_te(args) -> dict:
    from entropical_amd import Context, DeviceArray, combined_densities, datasets, te_autotune, te_matrix

    n_cols, n_rows, h, taus = 32, 100_000, 0.2, [1, 2, 5, 10]
    log("generating 32 coupled AR(1) series x 100000")
    X, _, _ = datasets.coupled_ar1(n_cols, n_rows)
    hs = [h] * n_cols
    ctx = Context(0)
    dx = DeviceArray(ctx, X)
    tun = te_autotune(ctx, dx, hs, taus) if not args.no_tune else ctx.get_tuning()
    log(f"tuning (q, m0, mr, group_h) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: te_matrix(ctx, dx, hs, taus), args.steps, args.warmup)
    pairs = sum(st["pairs_by_d"])
    log(f"TE (4 lags) {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {pairs:.3e} pairs")
    n_pairs = n_cols * (n_cols - 1) // 2
    jobs_3d, jobs_2d, jobs_1d = 2 * n_pairs * len(taus), n_pairs + n_cols * len(taus), n_cols
    dense = (jobs_3d + jobs_2d + jobs_1d) * float(n_rows) ** 2
    res = {
        "metric": "Gpair-evals/s (transfer entropy, 32 series x 100k frames, 4 lags)",
        "value": pairs / sec / 1e9, "unit": "Gpair-evals/s", "ms_per_step": 1e3 * sec,
        "config": {"workload": "te_32series_100000frames_h0.2_tau1-2-5-10",
                   "jobs": f"{jobs_3d} 3-D + {jobs_2d} 2-D + {jobs_1d} 1-D densities, one persistent launch",
                   "tuning_q_m0_mr_group": list(tun)},
        "pairs_candidate": pairs, "pairs_by_moments": st["moment_pairs"], "moment_terms": st["moment_terms"], "pairs_by_dim": st["pairs_by_d"], "lane_utilisation": util,
        "dense_equiv_gpair_per_s": dense / sec / 1e9,
        "roofline": roofline(st, kern_ms, "te"),
    }
    # the reference runs one lag per invocation: per lag 32 1-D + 32 lagged
    # 2-D + 496 2-D + 992 3-D combined_densities calls (transs.cpp:42-74)
    calls = {1: n_cols * len(taus), 2: (n_cols + n_pairs) * len(taus), 3: 2 * n_pairs * len(taus)}
    if not args.no_plain:
        log("timing plain per-pair path")
        tb = {}
        for d, cols, tau in ((1, [0], None), (2, [0, 1], None), (3, [0, 0, 1], [0, 5, 0])):
            hh = [h] * d
            combined_densities(X, n_rows, cols, hh, tau, method="brute", ctx=ctx)
            tb[d] = time_call(lambda: combined_densities(X, n_rows, cols, hh, tau, method="brute", ctx=ctx))
        total = sum(calls[d] * tb[d] for d in (1, 2, 3))
        res["plain_path"] = {"brute": {f"t_{d}d_ms": 1e3 * tb[d] for d in (1, 2, 3)} | {"te_ms": 1e3 * total}}
        res["speedup_vs_plain_brute"] = total / sec
    dx.free()
    ctx.close()
    if not args.no_cpu_baseline:
        th = args.cpu_threads or host_cores()
        tc = {1: cpu_omp_time(X[0:1], n_rows, [h], th), 2: cpu_omp_time(X[0:2], n_rows, [h, h], th)}
        lag = np.stack([X[0, :n_rows - 5], X[0, 5:], X[1, :n_rows - 5]])
        tc[3] = cpu_omp_time(lag, n_rows - 5, [h, h, h], th)
        total = sum(calls[d] * tc[d] for d in (1, 2, 3))
        res["cpu_baseline"] = {
            "value": total, "unit": "s per 4-lag TE matrix (extrapolated)", "cores": th, "kind": "port",
            "sample": "one 1-D, 2-D and 3-D density at N=100000 ({:.2f}, {:.2f}, {:.2f} s), scaled by the "
                      "reference's per-lag call counts".format(tc[1], tc[2], tc[3]),
        }
        res["speedup_vs_cpu"] = total / sec
    return res


# ---------------------------------------------------------------- uneven
def run_uneven(args) -> dict:
    from entropical_amd import Context, batch_autotune, batch_entropy, datasets

    log("generating 500 uneven trajectories")
    buf, jobs = datasets.uneven_batch(500)
    ctx = Context(0)
    ptr = ctx.alloc(buf.nbytes)
    ctx.h2d(ptr, buf)
    tun = batch_autotune(ctx, ptr, jobs) if not args.no_tune else ctx.get_tuning()
    log(f"tuning (q, m0, mr, group_h) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: batch_entropy(ctx, ptr, jobs), args.steps, args.warmup)
    pairs = sum(st["pairs_by_d"])
    dense = float(sum(j["n"] ** 2 for j in jobs))
    log(f"uneven batch {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {pairs:.3e} pairs")
    ctx.free(ptr)
    ctx.close()
    return {
        "metric": "Gpair-evals/s (500 uneven trajectories, densities + entropies)",
        "value": pairs / sec / 1e9, "unit": "Gpair-evals/s", "ms_per_step": 1e3 * sec,
        "config": {"workload": "uneven_500traj_2k-500k_d1-3", "points": int(sum(j["n"] for j in jobs)),
                   "tuning_q_m0_mr_group": list(tun)},
        "pairs_candidate": pairs, "pairs_by_moments": st["moment_pairs"], "moment_terms": st["moment_terms"], "pairs_by_dim": st["pairs_by_d"], "lane_utilisation": util,
        "dense_equiv_gpair_per_s": dense / sec / 1e9,
        "roofline": roofline(st, kern_ms, "uneven"),
    }


# ------------------------------------------------------------- entropy1d
def run_entropy1d(args) -> dict:
    from entropical_amd import Context, batch_autotune, batch_entropy, combined_densities, datasets

    n, h = 100_000, 0.1
    x = datasets.gmm_1d(n)
    ctx = Context(0)
    ptr = ctx.alloc(x.nbytes)
    ctx.h2d(ptr, x)
    spec = [{"n_dim": 1, "n": n, "n_rows": n, "off": [0], "h": [h]}]
    tun = batch_autotune(ctx, ptr, spec) if not args.no_tune else ctx.get_tuning()
    sec, kern_ms, st, util = timed_steps(ctx, lambda: batch_entropy(ctx, ptr, spec), args.steps, args.warmup)
    H = float(batch_entropy(ctx, ptr, spec)[0])
    pairs = sum(st["pairs_by_d"])
    res = {
        "metric": "Gpair-evals/s (1-D KDE entropy, N=100k)", "value": pairs / sec / 1e9, "unit": "Gpair-evals/s",
        "ms_per_step": 1e3 * sec,
        "config": {"workload": "entropy1d_gmm3_100000_h0.1", "tuning_q_m0_mr_group": list(tun)},
        "entropy": H, "pairs_candidate": pairs, "pairs_by_moments": st["moment_pairs"], "moment_terms": st["moment_terms"], "pairs_by_dim": st["pairs_by_d"], "lane_utilisation": util,
        "dense_equiv_gpair_per_s": float(n) ** 2 / sec / 1e9, "roofline": roofline(st, kern_ms, "entropy1d"),
    }
    if not args.no_plain:
        combined_densities(x, n, [0], [h], method="brute", ctx=ctx)
        tb = time_call(lambda: combined_densities(x, n, [0], [h], method="brute", ctx=ctx))
        res["plain_path"] = {"brute": {"t_1d_ms": 1e3 * tb}}
        res["speedup_vs_plain_brute"] = tb / sec
    ctx.free(ptr)
    ctx.close()
    if not args.no_cpu_baseline:
        th = args.cpu_threads or host_cores()
        t = cpu_omp_time(x, n, [h], th)
        res["cpu_baseline"] = {"value": t, "unit": "s per density", "cores": th, "kind": "port",
                               "sample": "the full workload (one 1-D density, N=100000)"}
        res["speedup_vs_cpu"] = t / sec
    return res


WORKLOADS = {"mi": run_mi, "te": run_te, "uneven": run_uneven, "entropy1d": run_entropy1d}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="mi")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a copy kernel of known traffic (counter calibration) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        print(json.dumps(copy_calibration()), flush=True)
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py measures one GPU only: this build has no multi-GPU path "
            "(jobs would be sharded by cost across GPUs with one input broadcast "
            "and one gather of per-job scalars); run with --gpus 1.",
            file=sys.stderr,
        )
        return 2
    res = WORKLOADS[args.workload](args)
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
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": res.pop("config"),
    }
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
