"""Seeded synthetic inputs for the benchmark workloads.

All arrays are column-major float32 ``(n_cols, n_rows)`` as
Tools::IO::read_coords(..., 'C') produces them (tools.hxx:226-228).
Generators use numpy's PCG64 (``numpy.random.default_rng``) with fixed seeds.
"""
from __future__ import annotations

import numpy as np

SEED_1D = 0xE17A0001
SEED_MI = 0xE17A0002
SEED_TE = 0xE17A0003
SEED_UNEVEN = 0xE17A0004
SEED_LEARN = 0xE17A0005


def gmm_1d(n: int = 100_000, seed: int = SEED_1D) -> np.ndarray:
    """3-component mixture: weights .3/.5/.2, means -1.5/0/2.0, sd .3/.5/.4."""
    rng = np.random.default_rng(seed)
    w = np.array([0.3, 0.5, 0.2])
    mu = np.array([-1.5, 0.0, 2.0])
    sd = np.array([0.3, 0.5, 0.4])
    comp = rng.choice(3, size=n, p=w)
    x = rng.standard_normal(n) * sd[comp] + mu[comp]
    return x.astype(np.float32).reshape(1, n)


def gmm_1d_pdf(x: np.ndarray) -> np.ndarray:
    w = np.array([0.3, 0.5, 0.2])
    mu = np.array([-1.5, 0.0, 2.0])
    sd = np.array([0.3, 0.5, 0.4])
    x = np.asarray(x, dtype=np.float64)[..., None]
    return (w * np.exp(-0.5 * ((x - mu) / sd) ** 2) / (sd * np.sqrt(2 * np.pi))).sum(-1)


def mixed_columns(n_cols: int = 64, n_rows: int = 200_000, seed: int = SEED_MI) -> np.ndarray:
    """Seeded n_cols x n_cols N(0,1) mixing of independent 2-component
    Gaussian mixtures, each column z-scored in FP64, returned as float32."""
    rng = np.random.default_rng(seed)
    S = np.empty((n_cols, n_rows), dtype=np.float64)
    for k in range(n_cols):
        w = rng.uniform(0.2, 0.8)
        m1, m2 = rng.uniform(-2.0, 2.0, size=2)
        s1, s2 = rng.uniform(0.3, 1.0, size=2)
        comp = rng.random(n_rows) < w
        S[k] = np.where(comp, m1 + s1 * rng.standard_normal(n_rows), m2 + s2 * rng.standard_normal(n_rows))
    A = rng.standard_normal((n_cols, n_cols))
    X = A @ S
    X -= X.mean(axis=1, keepdims=True)
    X /= X.std(axis=1, keepdims=True)
    return np.ascontiguousarray(X.astype(np.float32))


def coupled_ar1(n_series: int = 32, n_rows: int = 100_000, seed: int = SEED_TE):
    """Coupled AR(1) series x_{i,t} = 0.8 x_{i,t-1} + c_i x_{d(i),t-1} + e,
    e ~ N(0, sd 0.5), c_i ~ U(0, 0.3).  Series i >= 1 is driven by one
    earlier series d(i) < i drawn from the seed (series 0 is undriven), so
    the coupling matrix is triangular with spectral radius 0.8 (all-to-all
    coupling with 31 drivers of up to 0.3 each would diverge).
    Returns (coords (n_series, n_rows) float32, drivers, couplings)."""
    rng = np.random.default_rng(seed)
    drv = np.array([-1] + [int(rng.integers(0, i)) for i in range(1, n_series)])
    c = np.concatenate([[0.0], rng.uniform(0.0, 0.3, n_series - 1)])
    burn = 1000
    e = rng.standard_normal((n_series, n_rows + burn)) * 0.5
    x = np.zeros((n_series, n_rows + burn))
    for t in range(1, n_rows + burn):
        prev = x[:, t - 1]
        x[:, t] = 0.8 * prev + e[:, t]
        x[1:, t] += c[1:] * prev[drv[1:]]
    return np.ascontiguousarray(x[:, burn:].astype(np.float32)), drv, c


def uneven_batch(n_traj: int = 500, seed: int = SEED_UNEVEN, lmin: int = 2_000, lmax: int = 500_000):
    """Independent trajectories: length log-uniform in [lmin, lmax], dimension
    uniform in {1,2,3}, values = Student-t(nu=4) * 0.5 + a 2-component
    Gaussian mixture (means -1/+1, sd 0.3, equal weights), drawn
    independently per coordinate; bandwidth U(0.05, 0.3) per trajectory.
    Returns (flat float32 buffer, list of job dicts with offsets into it)."""
    rng = np.random.default_rng(seed)
    lens = np.rint(np.exp(rng.uniform(np.log(lmin), np.log(lmax), n_traj))).astype(np.int64)
    dims = rng.integers(1, 4, n_traj)
    hs = rng.uniform(0.05, 0.3, n_traj).astype(np.float32)
    total = int(np.sum(lens * dims))
    buf = np.empty(total, dtype=np.float32)
    jobs, off = [], 0
    for L, d, h in zip(lens, dims, hs):
        L, d = int(L), int(d)
        t = rng.standard_t(4, size=(d, L)) * 0.5
        g = np.where(rng.random((d, L)) < 0.5, -1.0, 1.0) + 0.3 * rng.standard_normal((d, L))
        buf[off: off + d * L] = (t + g).astype(np.float32).reshape(-1)
        jobs.append({"n_dim": d, "n": L, "n_rows": L, "off": [off + e * L for e in range(d)], "h": [float(h)] * d})
        off += d * L
    return buf, jobs
