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

Printed JSON line: value = pair evaluations per second (Gpair-evals/s) over
the whole step, counting the (query, candidate) pairs evaluated one by one.
Beside it: the candidate rate (every pair of every tile's cell-list window,
including the pairs of 1-D boxes and 2-D cells inside the support of every
query of a tile, which are summed exactly from their moments), the
dense-equivalent rate (sum of N^2), the plain per-pair GPU path (the
reference's per-call driver order, extrapolated from timed calls) and the
host OpenMP baseline (the oracle's restatement of densities_omp.cpp on a
stratified sample of the reference's calls, extrapolated).  The host leg
also checks sampled results against the FP64 oracle (max_rel_err_vs_fp64).
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


# The GPU box grants one GPU's share of its host: 16 cores (its
# OMP_NUM_THREADS), whatever the affinity mask shows.
BOX_CPU_SHARE = 16


def host_cores():
    """(threads used, cores in the affinity mask, reason)."""
    try:
        mask = len(os.sched_getaffinity(0))
    except AttributeError:
        mask = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or BOX_CPU_SHARE
    n = max(1, min(mask, share))
    why = (f"{n} threads: the host share granted with one GPU (OMP_NUM_THREADS={share}); "
           f"the affinity mask holds {mask} cores")
    return n, mask, why


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_oracle(threads: int):
    """The oracle (test infrastructure) for the host leg: OpenMP threads bound
    close (OMP_PROC_BIND is read when the runtime starts, at first load)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle

    oracle.set_threads(threads)
    return oracle


def rel(a: float, b: float) -> float:
    return abs(a - b) / abs(b)


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


def cpu_omp_time(orc, packed: np.ndarray, n: int, h) -> float:
    return time_call(lambda: orc.density_omp_packed(packed, n, h), reps=1)


def cpu_baseline_info(threads, mask, why, kind_note):
    return {"cores": threads, "kind": "port", "cores_note": why, "cpu_model": cpu_model(),
            "affinity_cores": mask, "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "method": kind_note}


def rates(st, sec: float) -> dict:
    """Pair counts of the last run and their rates over one step."""
    ev = sum(evaluated_pairs(st))
    cand = sum(st["pairs_by_d"])
    return {"pairs_evaluated": ev, "pairs_candidate": cand, "pairs_by_moments": st["moment_pairs"],
            "moment_terms": st["moment_terms"], "pairs_by_dim": st["pairs_by_d"],
            "pairs_evaluated_by_dim": evaluated_pairs(st), "candidate_gpair_per_s": cand / sec / 1e9}


def support_fields(r: dict, in_support: int, sec: float) -> dict:
    """Pairs inside the kernel's support (ordered (query, point) pairs with a
    nonzero term, the self pair included; exact count on the host): the work
    every exact method accounts for, whatever it evaluates.  The pairs summed
    from moments all lie inside it, so the in-support share of the pairs
    evaluated one by one is (in support - by moments) / evaluated."""
    return {
        "pairs_in_support": in_support,
        "in_support_gpair_per_s": in_support / sec / 1e9,
        "in_support_of_candidates": in_support / r["pairs_candidate"],
        "in_support_of_evaluated": (in_support - r["pairs_by_moments"]) / r["pairs_evaluated"],
        "method": "orc_count_in_support: sweep over dim 0, Fenwick tree over dim-1 ranks, |dx_e| < h_e in FP64",
    }


# ------------------------------------------------------------------ MI
def mi_call_sequence(n_cols: int) -> list:
    """The columns of every combined_densities call of Mi::main, in its order
    (mi.cpp:37-52): p_i, then for every j > i p_j and p_ij."""
    seq = []
    for i in range(n_cols):
        seq.append((i,))
        for j in range(i + 1, n_cols):
            seq.append((j,))
            seq.append((i, j))
    return seq


def stratified(seq: list, k: int) -> list:
    """k calls spread evenly over a call sequence."""
    return [seq[(2 * m + 1) * len(seq) // (2 * k)] for m in range(k)] if k > 0 else []


def run_mi(args) -> dict:
    from entropical_amd import Context, DeviceArray, batch_entropy, combined_densities, datasets, mi_autotune, mi_matrix

    n_cols, n_rows, h = 64, 200_000, 0.15
    log("generating 64 x 200000 mixed columns")
    X = datasets.mixed_columns(n_cols, n_rows)
    hs = [h] * n_cols
    ctx = Context(0)
    dx = DeviceArray(ctx, X)
    tun = mi_autotune(ctx, dx, hs) if not args.no_tune else ctx.get_tuning()
    log(f"tuning (q, m0, mr, group_h, sub) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: mi_matrix(ctx, dx, hs), args.steps, args.warmup)
    mi = mi_matrix(ctx, dx, hs)
    assert np.all(np.isfinite(mi)) and np.allclose(mi, mi.T)
    r = rates(st, sec)
    log(f"MI matrix {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {r['pairs_evaluated']:.3e} pairs "
        f"evaluated ({r['pairs_candidate']:.3e} candidates)")
    res = {
        "metric": "Gpair-evals/s (pairs evaluated one by one; batched 64-column MI matrix, 200k frames)",
        "value": r["pairs_evaluated"] / sec / 1e9,
        "unit": "Gpair-evals/s",
        "ms_per_step": 1e3 * sec,
        "config": {
            "workload": "mi_matrix_64col_200000frames_h0.15",
            "jobs": "64 1-D + 2016 2-D densities, one persistent launch",
            "n_cols": n_cols, "n_rows": n_rows, "bandwidth": h, "method": "cell list",
            "tuning_q_m0_mr_group_sub": list(tun),
        },
        **r,
        "lane_utilisation": util,
        "dense_equiv_gpair_per_s": 2080 * n_rows * n_rows / sec / 1e9,
        "roofline": roofline(st, kern_ms, "mi"),
    }
    # results the host leg checks against FP64 (outside the timed region)
    check_pairs = ((0, 1), (17, 42), (5, 63), (30, 31))
    H0 = None
    if not args.no_cpu_baseline:
        H0 = float(batch_entropy(ctx, dx.ptr, [{"n_dim": 1, "n": n_rows, "n_rows": n_rows, "off": [0],
                                                "h": [h]}])[0])
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
        th, mask, why = host_cores()
        orc = host_oracle(th)
        seq = mi_call_sequence(n_cols)
        one = [c for c in seq if len(c) == 1]
        two = [c for c in seq if len(c) == 2]
        k1, k2 = args.cpu_calls
        t1 = [cpu_omp_time(orc, X[list(c)], n_rows, [h]) for c in stratified(one, k1)]
        log(f"cpu 1-D densities {np.mean(t1):.2f} s each ({k1} calls)")
        t2 = [cpu_omp_time(orc, X[list(c)], n_rows, [h, h]) for c in stratified(two, k2)]
        log(f"cpu 2-D densities {np.mean(t2):.2f} s each ({k2} calls)")
        total = len(one) * float(np.mean(t1)) + len(two) * float(np.mean(t2))
        res["cpu_baseline"] = {
            "value": total, "unit": "s per MI matrix (extrapolated)",
            **cpu_baseline_info(th, mask, why, "oracle restatement of densities_omp.cpp (sorted-window scan, "
                                "OpenMP rows), -O3 -fopenmp -ftree-vectorize"),
            "sample": f"{k1} of mi.cpp's {len(one)} 1-D and {k2} of its {len(two)} 2-D combined_densities calls "
                      f"at N={n_rows}, spread evenly over its call order (mean {np.mean(t1):.2f} s and "
                      f"{np.mean(t2):.2f} s per call), extrapolated to all 4096 calls",
        }
        res["speedup_vs_cpu"] = total / sec
        log("pairs inside the support (exact count on the host)")
        pairs2 = [(i, j) for i in range(n_cols) for j in range(i + 1, n_cols)]
        ins = orc.count_in_support_jobs(X, n_rows, [[c] for c in range(n_cols)], [[h]] * n_cols) + \
            orc.count_in_support_jobs(X, n_rows, pairs2, [[h, h]] * len(pairs2))
        res["support"] = support_fields(r, ins, sec)
        log("FP64 checks of sampled results")
        # four fixed entries and the matrix's smallest |MI| (the worst
        # conditioned: its terms cancel the most)
        iu = np.triu_indices(n_cols, 1)
        k = int(np.argmin(np.abs(mi[iu])))
        check_pairs = check_pairs + ((int(iu[0][k]), int(iu[1][k])),)
        cols = sorted({c for ij in check_pairs for c in ij} | {0})
        p = {c: orc.density_fp64_window(X, n_rows, [c], [h]) for c in cols}
        errs, cond = {}, {}
        for i, j in check_pairs:
            pij = orc.density_fp64_window(X, n_rows, [i, j], [h, h])
            ref = orc.mi_fp64(pij, p[i], p[j])
            r = pij / (p[i] * p[j])
            terms = np.where(r > 0, pij * np.log(np.where(r > 0, r, 1.0)), 0.0)
            errs[f"mi[{i},{j}]"] = rel(mi[i, j], ref)
            cond[f"mi[{i},{j}]"] = float(np.abs(terms).sum() / abs(ref))
        errs["entropy[col 0]"] = rel(H0, orc.entropy(p[0]))
        cond["entropy[col 0]"] = float(np.abs(np.log(p[0])).sum() / abs(np.log(p[0]).sum()))
        bound = {k: max(1e-4, 1e-6 * cond[k]) for k in errs}
        res["max_rel_err_vs_fp64"] = {
            "max": max(errs.values()), "bound": "max(1e-4, 1e-6 * condition) per result",
            "within_bound": all(errs[k] <= bound[k] for k in errs), "results": errs, "condition": cond,
            "note": "relative error per result; condition = sum |terms| / |result|: the densities are within about "
                    "1e-6 of FP64, so a result's relative error grows with its condition (the smallest |MI| of the "
                    "matrix, mi[%d,%d], is the worst conditioned)" % check_pairs[-1]}
    return res


# ------------------------------------------------------------------ TE
def run_te(args) -> dict:
    from entropical_amd import Context, DeviceArray, combined_densities, datasets, te_autotune, te_matrix

    n_cols, n_rows, h, taus = 32, 100_000, 0.2, [1, 2, 5, 10]
    log("generating 32 coupled AR(1) series x 100000")
    X, drv, _ = datasets.coupled_ar1(n_cols, n_rows)
    hs = [h] * n_cols
    ctx = Context(0)
    dx = DeviceArray(ctx, X)
    tun = te_autotune(ctx, dx, hs, taus) if not args.no_tune else ctx.get_tuning()
    log(f"tuning (q, m0, mr, group_h, sub) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: te_matrix(ctx, dx, hs, taus), args.steps, args.warmup)
    te = te_matrix(ctx, dx, hs, taus)
    r = rates(st, sec)
    log(f"TE (4 lags) {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {r['pairs_evaluated']:.3e} pairs")
    n_pairs = n_cols * (n_cols - 1) // 2
    jobs_3d, jobs_2d, jobs_1d = 2 * n_pairs * len(taus), n_pairs + n_cols * len(taus), n_cols
    dense = (jobs_3d + jobs_2d + jobs_1d) * float(n_rows) ** 2
    res = {
        "metric": "Gpair-evals/s (pairs evaluated one by one; transfer entropy, 32 series x 100k frames, 4 lags)",
        "value": r["pairs_evaluated"] / sec / 1e9, "unit": "Gpair-evals/s", "ms_per_step": 1e3 * sec,
        "config": {"workload": "te_32series_100000frames_h0.2_tau1-2-5-10",
                   "jobs": f"{jobs_3d} 3-D + {jobs_2d} 2-D + {jobs_1d} 1-D densities, one persistent launch",
                   "tuning_q_m0_mr_group_sub": list(tun)},
        **r, "lane_utilisation": util,
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
        th, mask, why = host_cores()
        orc = host_oracle(th)
        tc = {1: cpu_omp_time(orc, X[[0]], n_rows, [h]), 2: cpu_omp_time(orc, X[[0, 1]], n_rows, [h, h])}
        lag = np.stack([X[0, :n_rows - 5], X[0, 5:], X[1, :n_rows - 5]])
        tc[3] = cpu_omp_time(orc, lag, n_rows - 5, [h, h, h])
        total = sum(calls[d] * tc[d] for d in (1, 2, 3))
        res["cpu_baseline"] = {
            "value": total, "unit": "s per 4-lag TE matrix (extrapolated)",
            **cpu_baseline_info(th, mask, why, "oracle restatement of densities_omp.cpp (packed columns, lag "
                                "applied), -O3 -fopenmp -ftree-vectorize"),
            "sample": "one 1-D, 2-D and 3-D density at N=100000 ({:.2f}, {:.2f}, {:.2f} s), scaled by the "
                      "reference's per-lag call counts".format(tc[1], tc[2], tc[3]),
        }
        res["speedup_vs_cpu"] = total / sec
        log("FP64 checks of sampled results")
        i = 5
        j = int(drv[i])
        a, b = min(i, j), max(i, j)
        tau = taus[0]
        m = n_rows - tau
        f = lambda v: float(np.float32(v))
        pa = orc.density_fp64_window(X, n_rows, [a], [h])
        pb = orc.density_fp64_window(X, n_rows, [b], [h])
        pab = orc.density_fp64_window(X, n_rows, [a, b], [h, h])
        errs = {}
        for (u, v, pu, cols, lags) in ((a, b, pa, [a, a, b], [0, tau, 0]), (b, a, pb, [a, b, b], [0, 0, tau])):
            p3 = orc.density_fp64_window(X, n_rows, cols, [h, h, h], lags)
            ptau = orc.density_fp64_window(X, n_rows, [u, u], [h, h], [0, tau])
            ref = orc.te_fp64(p3, pu[:m], pab[:m], ptau, f(h) ** 3)
            errs[f"te[tau={tau}][{u},{v}]"] = rel(te[0, u, v], ref)
        res["max_rel_err_vs_fp64"] = {"max": max(errs.values()), "bound": 1e-4, "results": errs}
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
    log(f"tuning (q, m0, mr, group_h, sub) = {tun}")
    sec, kern_ms, st, util = timed_steps(ctx, lambda: batch_entropy(ctx, ptr, jobs), args.steps, args.warmup)
    H = batch_entropy(ctx, ptr, jobs)
    r = rates(st, sec)
    dense = float(sum(j["n"] ** 2 for j in jobs))
    log(f"uneven batch {1e3 * sec:.1f} ms/step, density kernel {kern_ms:.1f} ms, {r['pairs_evaluated']:.3e} pairs")
    ctx.free(ptr)
    ctx.close()
    res = {
        "metric": "Gpair-evals/s (pairs evaluated one by one; 500 uneven trajectories, densities + entropies)",
        "value": r["pairs_evaluated"] / sec / 1e9, "unit": "Gpair-evals/s", "ms_per_step": 1e3 * sec,
        "config": {"workload": "uneven_500traj_2k-500k_d1-3", "points": int(sum(j["n"] for j in jobs)),
                   "tuning_q_m0_mr_group_sub": list(tun)},
        **r, "lane_utilisation": util,
        "dense_equiv_gpair_per_s": dense / sec / 1e9,
        "roofline": roofline(st, kern_ms, "uneven"),
    }
    if not args.no_cpu_baseline:
        th, mask, why = host_cores()
        orc = host_oracle(th)
        log("FP64 checks of sampled results")
        errs = {}
        for d in (1, 2):
            k = max((s for s in range(len(jobs)) if jobs[s]["n_dim"] == d), key=lambda s: jobs[s]["n"])
            s = jobs[k]
            pts = np.stack([buf[s["off"][e]: s["off"][e] + s["n"]] for e in range(d)])
            t0 = time.perf_counter()
            orc.density_omp_packed(pts, s["n"], s["h"][:d])
            errs[f"cpu_{d}d_s"] = time.perf_counter() - t0
            p = orc.density_fp64_window(pts, s["n"], list(range(d)), s["h"][:d])
            errs[f"entropy[traj {k}, d={d}, n={s['n']}]"] = rel(H[k], orc.entropy(p))
        t1, t2 = errs.pop("cpu_1d_s"), errs.pop("cpu_2d_s")
        res["max_rel_err_vs_fp64"] = {"max": max(errs.values()), "bound": 1e-4, "results": errs}
        res["cpu_baseline"] = {
            "value": t1 + t2, "unit": "s for the longest 1-D and 2-D trajectories",
            **cpu_baseline_info(th, mask, why, "oracle restatement of densities_omp.cpp"),
            "sample": "the batch's longest 1-D and 2-D trajectory ({:.2f} s, {:.2f} s); the whole batch is "
                      "not extrapolated (its cost is not linear in the sample)".format(t1, t2),
        }
    return res


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
    dens = combined_densities(x, n, [0], [h], ctx=ctx) if not args.no_cpu_baseline else None
    r = rates(st, sec)
    res = {
        "metric": "Gpair-evals/s (pairs evaluated one by one; 1-D KDE entropy, N=100k)",
        "value": r["pairs_evaluated"] / sec / 1e9, "unit": "Gpair-evals/s",
        "ms_per_step": 1e3 * sec,
        "config": {"workload": "entropy1d_gmm3_100000_h0.1", "tuning_q_m0_mr_group_sub": list(tun)},
        "entropy": H, **r, "lane_utilisation": util,
        "dense_equiv_gpair_per_s": float(n) ** 2 / sec / 1e9, "roofline": roofline(st, kern_ms, "entropy1d"),
    }
    if not args.no_plain:
        combined_densities(x, n, [0], [h], method="brute", ctx=ctx)
        tb = time_call(lambda: combined_densities(x, n, [0], [h], method="brute", ctx=ctx))
        tc = time_call(lambda: combined_densities(x, n, [0], [h], method="cell", ctx=ctx))
        res["plain_path"] = {"brute": {"t_1d_ms": 1e3 * tb}, "cell": {"t_1d_ms": 1e3 * tc}}
        res["speedup_vs_plain_brute"] = tb / sec
        res["speedup_vs_plain_cell"] = tc / sec
    ctx.free(ptr)
    ctx.close()
    if not args.no_cpu_baseline:
        th, mask, why = host_cores()
        orc = host_oracle(th)
        t = cpu_omp_time(orc, x, n, [h])
        res["cpu_baseline"] = {"value": t, "unit": "s per density",
                               **cpu_baseline_info(th, mask, why, "oracle restatement of densities_omp.cpp"),
                               "sample": "the full workload (one 1-D density, N=100000)"}
        res["speedup_vs_cpu"] = t / sec
        res["support"] = support_fields(r, orc.count_in_support(x, n, [0], [h]), sec)
        ref = orc.density_fp64_window(x, n, [0], [h])
        errs = {"density (max over 100000 points)": float(np.max(np.abs(dens - ref) / ref)),
                "entropy": rel(H, orc.entropy(ref))}
        res["max_rel_err_vs_fp64"] = {"max": errs["entropy"], "bound": 1e-4, "results": errs,
                                      "density_bound": 1e-5}
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
    ap.add_argument("--cpu-calls", type=lambda v: tuple(int(a) for a in v.split(",")), default=(4, 12),
                    help="MI host baseline: 1-D,2-D calls timed (16,64 is the sample of the measurement plan)")
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
