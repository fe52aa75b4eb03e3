"""Randomized parity: seeded random jobs (1-3 dims, lags, data shapes from
clustered to heavy-tailed, coordinate scales and offsets, tile shapes, box
and cell moments on and off) through the C ABI against the FP64 oracle at the
density tolerance, 1e-5 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5
TUNINGS = [(1, 32, 2, 64), (2, 32, 1, 64), (1, 16, 2, 16), (4, 8, 1, 4), (2, 64, 2, 8), (1, 12, 3, 5),
           (1, 32, 4, 64, 16), (1, 16, 4, 16, 32)]


def sample(r, n):
    kind = r.choice(["normal", "mix", "t", "ties", "cluster", "ar"])
    if kind == "normal":
        x = r.standard_normal((3, n))
    elif kind == "mix":
        x = np.where(r.random((3, n)) < 0.3, r.normal(-2, 0.3, (3, n)), r.normal(1, 0.7, (3, n)))
    elif kind == "t":
        x = r.standard_t(float(r.choice([1.5, 3.0])), (3, n))
    elif kind == "ties":
        x = np.round(r.standard_normal((3, n)) * 10) / 10
    elif kind == "cluster":
        x = r.normal(0, 0.01, (3, n)) + r.integers(0, 3, (3, n))
    else:
        x = np.zeros((3, n))
        e = r.normal(0, 0.5, (3, n))
        for t in range(1, n):
            x[:, t] = 0.8 * x[:, t - 1] + e[:, t]
    return x, kind


def test_random_jobs(ctx, orc, monkeypatch):
    from entropical_amd import combined_densities

    r = np.random.default_rng(20261203)
    failures = []
    try:
        for case in range(64):
            d = int(r.integers(1, 4))
            n = int(r.choice([300, 3000, 20000]))
            x, kind = sample(r, n)
            scale = float(r.choice([1e-2, 1.0, 30.0]))
            off = float(r.choice([0.0, 50.0, -3e3]))
            x = (x * scale + off).astype(np.float32)
            cols = [int(c) for c in r.integers(0, 3, size=d)]
            tau = [int(t) for t in r.integers(0, 6, size=d)] if r.random() < 0.3 else None
            h = [float(v) * scale for v in r.uniform(0.02, 0.5, size=d)]
            tuning = TUNINGS[case % len(TUNINGS)]
            monkeypatch.setenv("ENT_BOX_MOMENTS", "0" if case % 4 == 3 else "1")
            ctx.set_tuning(*tuning)
            got = combined_densities(x, n, cols, h, tau, ctx=ctx)
            ref = orc.density_fp64_window(x, n, cols, h, tau)
            err = float(np.max(np.abs(got - ref) / np.abs(ref)))
            if not err < RTOL:
                failures.append((case, d, n, kind, scale, off, cols, tau, h, tuning, err))
    finally:
        ctx.set_tuning()
    assert not failures, failures
