"""GPU tests of the tile-shape autotuner (ent_mi_autotune_dev,
ent_autotune_dev): the tuning it returns is a member of its grid and is
left on the context, densities computed with it match FP64, the
ENT_TUNE_CACHE file round trip, the workload signature's bandwidth/range
term, and the caller's tuning restored when no grid point supports a
workload.  The autotuner replaces the reference's compile-time BSIZE
(densities_cuda.cu:14)."""
import numpy as np
import pytest

from entropical_amd import (Context, DeviceArray, EntropicalError, batch_autotune, combined_densities,
                            mi_autotune, mi_matrix)

pytestmark = pytest.mark.gpu

GRID_Q, GRID_M0, GRID_MR, GRID_GH = (1, 2, 4), (8, 16, 32), (1, 2), (4, 16, 64)
# sub-tiles of 16: q = 1, group_h = 64, and (m0, mr) from the 2-D grid
# (batches with 2-D jobs) or the 3-D one (batches with 3-D jobs)
SUB_Q, SUB_M0, SUB_MR = (1,), (32, 64), (2, 3, 4)
SUB3_M0, SUB3_MR = (16, 32), (1, 2)


def in_grid(t, has_2d=True, has_3d=False):
    q, m0, mr, gh, sub = t
    if sub == 0:
        return q in GRID_Q and m0 in GRID_M0 and mr in GRID_MR and gh in GRID_GH
    if sub != 16 or q not in SUB_Q or gh != 64:
        return False
    return (has_2d and m0 in SUB_M0 and mr in SUB_MR) or (has_3d and m0 in SUB3_M0 and mr in SUB3_MR)


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def test_mi_autotune_grid_member_parity_and_cache(orc, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("ENT_TUNE_CACHE", str(cache))
    r = np.random.default_rng(201)
    n, c = 20000, 6
    a = r.standard_normal((c, n))
    a[1] += 0.7 * a[0]
    a[3] = np.tanh(a[2]) + 0.3 * a[3]
    x = a.astype(np.float32)
    h = [0.2, 0.25, 0.15, 0.3, 0.2, 0.22]
    with Context(0) as ctx:
        dx = DeviceArray(ctx, x)
        try:
            t = mi_autotune(ctx, dx, h)
            assert in_grid(t), t
            assert ctx.get_tuning() == t
            lines = cache.read_text().splitlines()
            assert len(lines) == 1 and lines[0].split()[1:] == [str(v) for v in t]
            # densities and MI at the chosen tuning
            for cols in ([0], [0, 1], [2, 3]):
                hh = [h[i] for i in cols]
                got = combined_densities(x, n, cols, hh, ctx=ctx)
                assert rel_err(got, orc.density_fp64_window(x, n, cols, hh)) < 1e-5, cols
            mi = mi_matrix(ctx, dx, h)
            p = [orc.density_fp64_window(x, n, [i], [h[i]]) for i in (0, 1)]
            ref = orc.mi_fp64(orc.density_fp64_window(x, n, [0, 1], [h[0], h[1]]), p[0], p[1])
            assert mi[0, 1] == pytest.approx(ref, rel=1e-4)
        finally:
            dx.free()
    # a fresh context reads the cache file: same tuning, nothing appended
    with Context(0) as ctx2:
        dx = DeviceArray(ctx2, x)
        try:
            ctx2.set_tuning(1, 8, 1, 4)
            assert mi_autotune(ctx2, dx, h) == t
            assert ctx2.get_tuning() == t
        finally:
            dx.free()
    assert len(cache.read_text().splitlines()) == 1


def test_batch_autotune_signature_and_cache_hit(ctx, orc, monkeypatch, tmp_path):
    monkeypatch.setenv("ENT_TUNE_CACHE", str(tmp_path / "t.txt"))
    r = np.random.default_rng(202)
    n = 15000
    x = r.standard_normal((3, n)).astype(np.float32)
    dx = DeviceArray(ctx, x)
    try:
        def jobs(scale):
            return [{"n_dim": 1, "n": n, "off": [0], "h": [0.1 * scale]},
                    {"n_dim": 2, "n": n, "off": [0, n], "h": [0.15 * scale, 0.2 * scale]},
                    {"n_dim": 3, "n": n, "off": [0, n, 2 * n], "h": [0.3 * scale, 0.3 * scale, 0.3 * scale]}]
        t1, ms1 = batch_autotune(ctx, dx.ptr, jobs(1.0), return_ms=True)
        assert in_grid(t1, has_3d=True) and ms1 > 0
        t2, ms2 = batch_autotune(ctx, dx.ptr, jobs(1.0), return_ms=True)
        assert t2 == t1 and ms2 == -1.0  # cache hit
        # the same jobs at 4x the bandwidths: another signature, searched again
        t3, ms3 = batch_autotune(ctx, dx.ptr, jobs(4.0), return_ms=True)
        assert in_grid(t3, has_3d=True) and ms3 > 0
        # densities at the tuning left on the context
        got = combined_densities(x, n, [0, 1, 2], [1.2, 1.2, 1.2], ctx=ctx)
        assert rel_err(got, orc.density_fp64_window(x, n, [0, 1, 2], [1.2, 1.2, 1.2])) < 1e-5
    finally:
        dx.free()
        ctx.set_tuning()


def test_autotune_skips_unsupported_shapes_and_restores_on_failure(ctx, orc):
    r = np.random.default_rng(203)
    n = 2000
    # range / h = 1e6: 8 or 16 dim-0 boxes per bandwidth stay below 2^24
    # boxes, 32 do not -- those grid points are skipped
    x = r.uniform(0.0, 1.0e6, size=(1, n)).astype(np.float32)
    x[0, 0], x[0, 1] = 0.0, 1.0e6
    dx = DeviceArray(ctx, x)
    try:
        ctx.set_tuning(2, 8, 1, 4)
        t = batch_autotune(ctx, dx.ptr, [{"n_dim": 1, "n": n, "off": [0], "h": [1.0]}])
        assert in_grid(t, has_2d=False) and t[1] in (8, 16) and t[4] == 0, t
        got = combined_densities(x, n, [0], [1.0], ctx=ctx)
        assert rel_err(got, orc.density_fp64(x, n, [0], [1.0])) < 1e-5
        # range / h = 1e7: no grid point fits; the error leaves the caller's tuning
        ctx.set_tuning(1, 16, 2, 16, 16)
        with pytest.raises(EntropicalError):
            batch_autotune(ctx, dx.ptr, [{"n_dim": 1, "n": n, "off": [0], "h": [0.1]}])
        assert ctx.get_tuning() == (1, 16, 2, 16, 16)
    finally:
        dx.free()
        ctx.set_tuning()
