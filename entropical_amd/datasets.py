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
