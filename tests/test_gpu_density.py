"""GPU parity: the HIP density path (through the C ABI) against the FP64
oracle.  Tolerance: densities within 1e-5 relative of FP64 (FP32 compensated
accumulation on the device, FP64 normalisation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def data(seed, n_cols, n, scale=1.0, kind="normal"):
    r = np.random.default_rng(seed)
    if kind == "normal":
        x = r.standard_normal((n_cols, n))
    elif kind == "t4":
        x = r.standard_t(4, size=(n_cols, n))
    elif kind == "uniform":
        x = r.uniform(-1, 1, size=(n_cols, n))
    else:
        raise ValueError(kind)
    return (scale * x).astype(np.float32)


@pytest.mark.parametrize("method", ["cell", "brute"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_density_matches_fp64(ctx, orc, method, d):
    from entropical_amd import combined_densities

    n = 3000
    x = data(100 + d, 3, n)
    cols = [0, 1, 2][:d]
    h = [0.2, 0.3, 0.45][:d]
    got = combined_densities(x, n, cols, h, method=method, ctx=ctx)
    ref = orc.density_fp64(x, n, cols, h)
    assert got.shape == ref.shape
    assert rel_err(got, ref) < RTOL


@pytest.mark.parametrize("method", ["cell", "brute"])
def test_lagged_and_repeated_columns(ctx, orc, method):
    from entropical_amd import combined_densities

    n = 2500
    x = data(7, 2, n)
    for cols, tau in (([0, 0], [0, 5]), ([0, 0, 1], [0, 3, 0]), ([0, 1, 1], [0, 0, 10]), ([1], [4])):
        h = [0.3, 0.25, 0.4][: len(cols)]
        got = combined_densities(x, n, cols, h, tau, method=method, ctx=ctx)
        ref = orc.density_fp64(x, n, cols, h, tau)
        assert got.shape == (n - max(tau),)
        assert rel_err(got, ref) < RTOL, (cols, tau)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 257, 4099])
def test_small_and_ragged_sizes(ctx, orc, n):
    from entropical_amd import combined_densities

    x = data(n, 3, n)
    for d in (1, 2, 3):
        cols = list(range(d))
        h = [0.5] * d
        ref = orc.density_fp64(x, n, cols, h)
        for method in ("cell", "brute"):
            got = combined_densities(x, n, cols, h, method=method, ctx=ctx)
            assert rel_err(got, ref) < RTOL, (d, method)


def test_heavy_tails_and_ties(ctx, orc):
    from entropical_amd import combined_densities

    n = 6000
    x = data(11, 3, n, kind="t4", scale=3.0)
    x[1, ::7] = 0.25  # many exact ties in one dim
    x[2, 100:200] = x[2, 100]  # a block of identical values
    for cols, h in (([0], [0.05]), ([0, 1], [0.1, 0.07]), ([0, 1, 2], [0.3, 0.2, 0.25])):
        got = combined_densities(x, n, cols, h, ctx=ctx)
        ref = orc.density_fp64_window(x, n, cols, h)
        assert rel_err(got, ref) < RTOL, cols


def test_all_points_identical(ctx):
    from entropical_amd import combined_densities

    n = 1000
    x = np.full((2, n), 1.5, np.float32)
    got = combined_densities(x, n, [0, 1], [0.2, 0.4], ctx=ctx)
    h1, h2 = float(np.float32(0.2)), float(np.float32(0.4))
    np.testing.assert_allclose(got, (0.75 / h1) * (0.75 / h2), rtol=1e-6)


@pytest.mark.parametrize("tuning", [(1, 1, 1, 1), (2, 2, 1, 4), (4, 8, 2, 1), (4, 4, 3, 2), (1, 32, 4, 64, 16),
                                    (1, 16, 2, 16, 32), (1, 8, 5, 2, 16)])
def test_tunings_agree(ctx, orc, tuning):
    from entropical_amd import combined_densities

    n = 5000
    x = data(21, 3, n)
    ctx.set_tuning(*tuning)
    try:
        for d in (1, 2, 3):
            cols = list(range(d))
            h = [0.15, 0.2, 0.3][:d]
            got = combined_densities(x, n, cols, h, ctx=ctx)
            ref = orc.density_fp64_window(x, n, cols, h)
            assert rel_err(got, ref) < RTOL, (tuning, d)
    finally:
        ctx.set_tuning()


def test_deterministic(ctx):
    from entropical_amd import combined_densities

    n = 20000
    x = data(31, 2, n)
    a = combined_densities(x, n, [0, 1], [0.1, 0.1], ctx=ctx)
    b = combined_densities(x, n, [0, 1], [0.1, 0.1], ctx=ctx)
    np.testing.assert_array_equal(a, b)


def test_nonfinite_is_rejected(ctx):
    from entropical_amd import EntropicalError, combined_densities

    x = data(1, 1, 100)
    x[0, 17] = np.nan
    with pytest.raises(EntropicalError, match="non-finite"):
        combined_densities(x, 100, [0], [0.3], ctx=ctx)


# ------------------------------------------------------------- binning
def test_bin_bit_exact_vs_oracle(ctx, orc):
    from entropical_amd import bin_column

    r = np.random.default_rng(5)
    for trial in range(30):
        n = int(r.integers(1, 50000))
        x = (r.standard_normal(n) * r.uniform(0.01, 50) + r.uniform(-100, 100)).astype(np.float32)
        bs = np.float32(r.uniform(1e-3, 2.0))
        g_box, g_nb, g_mn, g_mx = bin_column(x, bs, ctx=ctx)
        o_box, o_nb, o_mn, o_mx = orc.bin_column(x, bs)
        assert (g_nb, g_mn, g_mx) == (o_nb, o_mn, o_mx)
        np.testing.assert_array_equal(g_box, o_box)


def test_bin_edge_cases(ctx, orc):
    from entropical_amd import bin_column

    cases = [
        (np.array([0.0, 0.1, 0.2, 0.3, 0.95, 1.0], np.float32), 0.25),
        (np.full(5, -2.5, np.float32), 0.1),
        (np.array([-3.0, -2.0, -1.5], np.float32), 0.5),
        (np.arange(1000, dtype=np.float32) * np.float32(0.1), np.float32(0.1)),
    ]
    for x, bs in cases:
        g = bin_column(x, bs, ctx=ctx)
        o = orc.bin_column(x, bs)
        assert g[1:] == o[1:]
        np.testing.assert_array_equal(g[0], o[0])


@pytest.mark.parametrize("sort_mode,scatter", [("auto", "staged"), ("shared", "staged"), ("full", "direct")])
def test_cell_sort_bit_exact(ctx, orc, monkeypatch, sort_mode, scatter):
    """GPU stable cell sort == CPU stable counting sort by the same keys, for
    the one-sort-per-job and the shared dim-0 sort schemes and both scatter
    kernels."""
    from entropical_amd import cell_sort

    if sort_mode != "auto":
        monkeypatch.setenv("ENT_SORT", sort_mode)
    monkeypatch.setenv("ENT_SCATTER", scatter)
    r = np.random.default_rng(9)
    n = 30000
    x = r.standard_normal((3, n)).astype(np.float32)
    for cols, h, tau in (([0], [0.1], None), ([0, 1], [0.2, 0.3], None), ([0, 1, 2], [0.3, 0.2, 0.4], [0, 2, 1])):
        keys, perm, nb, mn, bs = cell_sort(x, n, cols, h, tau, ctx=ctx)
        d = len(cols)
        tau = tau or [0] * d
        ns = n - max(tau)
        # CPU: boxes per dim with the oracle's binning, same box sizes
        boxes = []
        for e in range(d):
            col = x[cols[e], tau[e]: tau[e] + ns]
            b, onb, omn, _ = orc.bin_column(col, bs[e])
            assert onb == nb[e] and omn == mn[e]
            boxes.append(b.astype(np.uint64))
        key = boxes[0].copy()
        if d >= 2:
            row = boxes[1].copy()
            if d == 3:
                row += boxes[2] * np.uint64(nb[1])
            key += row * np.uint64(nb[0])
        order = np.argsort(key, kind="stable")
        np.testing.assert_array_equal(perm, order.astype(np.uint32))
        np.testing.assert_array_equal(keys, key[order])


@pytest.mark.parametrize("d", [2, 3])
def test_packed_small_tiles(ctx, orc, monkeypatch, d):
    """Sparse tails give tiles of a few queries, which run packed two or four
    to a wave (16 or 32 lanes per tile).  Densities with packing and without
    it (ENT_TILE_PACK=0) both match FP64, and packing issues fewer lane
    slots, so the packed path ran."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(31 + d)
    n = 6000
    x = r.standard_t(3, size=(3, n)).astype(np.float32)
    cols = [0, 1, 2][:d]
    h = [0.1, 0.12, 0.15][:d]
    ref = orc.density_fp64(x, n, cols, h)
    monkeypatch.setenv("ENT_TILE_PACK", "0")
    single = combined_densities(x, n, cols, h, ctx=ctx)
    slots_single = ctx.last_stats()["lane_pairs"]
    monkeypatch.setenv("ENT_TILE_PACK", "1")
    packed = combined_densities(x, n, cols, h, ctx=ctx)
    slots_packed = ctx.last_stats()["lane_pairs"]
    assert rel_err(single, ref) <= RTOL
    assert rel_err(packed, ref) <= RTOL
    assert slots_packed < slots_single


@pytest.mark.gpu
@pytest.mark.parametrize("h", [0.05, 0.2])
def test_box_moments_1d(ctx, orc, monkeypatch, h):
    """1-D boxes inside the support of every query of a tile are summed from
    their moments (ENT_BOX_MOMENTS, on by default).  With moments and without
    them the densities match FP64, the candidate pair count is the same, and
    with moments on some of those pairs (and none with them off) come from
    box moments."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(41)
    n = 60000
    x = np.concatenate([r.standard_normal(n // 2), 0.7 * r.standard_t(4, size=n - n // 2)]).astype(np.float32)
    x = r.permutation(x)[None, :]
    ref = orc.density_fp64(x, n, [0], [h])
    monkeypatch.setenv("ENT_BOX_MOMENTS", "0")
    plain = combined_densities(x, n, [0], [h], ctx=ctx)
    st0 = ctx.last_stats()
    monkeypatch.setenv("ENT_BOX_MOMENTS", "1")
    mom = combined_densities(x, n, [0], [h], ctx=ctx)
    st1 = ctx.last_stats()
    assert rel_err(plain, ref) <= RTOL
    assert rel_err(mom, ref) <= RTOL
    assert st0["moment_pairs"] == 0 and st0["moment_terms"] == 0
    assert st1["pairs"] == st0["pairs"]
    assert 0 < st1["moment_pairs"] < st1["pairs"]
    assert 0 < st1["moment_terms"] < st1["moment_pairs"]


@pytest.mark.gpu
@pytest.mark.parametrize("h", [(0.1, 0.15), (0.3, 0.2)])
def test_cell_moments_2d(ctx, orc, monkeypatch, h):
    """2-D cells (one row high, about h/4 wide) inside the support of every
    query of a dense tile are summed from their 9 moments.  With moments and
    without them the densities match FP64 and the candidate pair count is
    the same; with moments some 2-D pairs come from cell moments."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(43)
    n = 80000
    a = r.standard_normal(n)
    x = np.stack([a, 0.6 * a + 0.8 * r.standard_normal(n)]).astype(np.float32)
    ref = orc.density_fp64(x, n, [0, 1], list(h))
    monkeypatch.setenv("ENT_BOX_MOMENTS", "0")
    plain = combined_densities(x, n, [0, 1], list(h), ctx=ctx)
    st0 = ctx.last_stats()
    monkeypatch.setenv("ENT_BOX_MOMENTS", "1")
    mom = combined_densities(x, n, [0, 1], list(h), ctx=ctx)
    st1 = ctx.last_stats()
    assert rel_err(plain, ref) <= RTOL
    assert rel_err(mom, ref) <= RTOL
    assert st0["moment_pairs"] == 0
    assert st1["pairs"] == st0["pairs"]
    assert 0 < st1["moment_pairs_by_d"][1] < st1["pairs"]
    assert 0 < st1["moment_terms_by_d"][1] < st1["moment_pairs_by_d"][1]


@pytest.mark.parametrize("sub,mr", [(16, 4), (32, 4), (16, 2)])
def test_sub_tiles_2d(ctx, orc, monkeypatch, sub, mr):
    """2-D tiles run as sub-tiles of 16 or 32 queries, each with its own
    bounding box, moment cells and candidate stream.  With moments on and
    off the densities match FP64 and equal the whole-tile path's within the
    tolerance; the sub-tiles' candidate windows are smaller, so fewer pairs
    are evaluated one by one than with whole tiles of the same shape; with
    moments on, some pairs come from cell moments."""
    from entropical_amd import EntropicalError, combined_densities

    with pytest.raises(EntropicalError, match="q = 1"):
        ctx.set_tuning(2, 32, mr, 64, sub)  # sub-tiles run with one query per lane
    r = np.random.default_rng(151 + sub + mr)
    n = 80000
    a = r.standard_normal(n)
    x = np.stack([a, 0.6 * a + 0.8 * r.standard_normal(n)]).astype(np.float32)
    # bandwidths for which the job's moment table (cells h/8 wide, rows h/mr
    # high) has fewer cells than points, so it is built
    h = [0.2, 0.25]
    ref = orc.density_fp64_window(x, n, [0, 1], h)
    got, st = {}, {}
    try:
        for s in (0, sub):
            for mom in ("0", "1"):
                monkeypatch.setenv("ENT_BOX_MOMENTS", mom)
                ctx.set_tuning(1, 32, mr, 64, s)
                got[s, mom] = combined_densities(x, n, [0, 1], h, ctx=ctx)
                st[s, mom] = ctx.last_stats()
                assert rel_err(got[s, mom], ref) < RTOL, (s, mom)
    finally:
        ctx.set_tuning()

    def evaluated(t):
        return t["pairs"] - t["moment_pairs"]

    # same candidate windows with moments on and off
    assert st[sub, "0"]["pairs"] == st[sub, "1"]["pairs"]
    assert st[sub, "0"]["moment_pairs"] == 0
    assert 0 < st[sub, "1"]["moment_pairs_by_d"][1] < st[sub, "1"]["pairs"]
    assert evaluated(st[sub, "0"]) < 0.9 * evaluated(st[0, "0"])
    assert evaluated(st[sub, "1"]) < 0.9 * evaluated(st[0, "1"])


def test_sub_tiles_packed_and_sparse(ctx, orc, monkeypatch):
    """Heavy tails give short 2-D tiles: those of at most 16 or 32 queries
    are packed (ENT_TILE_PACK) and run through the sub-tile path with the
    offsets of the pack, the others are split; tiles wider than 32
    bandwidths take raw-coordinate pair terms.  All match FP64 with packing
    on and off."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(157)
    n = 30000
    x = (r.standard_t(1.5, size=(2, n)) * 2.0).astype(np.float32)
    h = [0.05, 0.3]
    ref = orc.density_fp64_window(x, n, [0, 1], h)
    try:
        for sub in (16, 32):
            ctx.set_tuning(1, 32, 4, 64, sub)
            for pack in ("0", "1"):
                monkeypatch.setenv("ENT_TILE_PACK", pack)
                got = combined_densities(x, n, [0, 1], h, ctx=ctx)
                assert rel_err(got, ref) < RTOL, (sub, pack)
            assert ctx.last_stats()["raw_tiles"] > 0
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("sub,mr", [(16, 1), (32, 1), (16, 2)])
def test_sub_tiles_3d(ctx, orc, sub, mr):
    """3-D tiles run as sub-tiles of 16 or 32 queries (the one-query-per-lane
    kernel variant with 3-D sub-tiles), each with its own bounding box and
    candidate rows (two range entries per lane: up to 2W rows).  The
    densities match FP64 and the whole-tile path's within the tolerance; the
    sub-tiles' candidate windows are narrower along dim 0, so fewer pairs are
    evaluated than with whole tiles of the same shape."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(161 + sub + mr)
    n = 60000
    # (x_t, x_t+1, y_t) of an AR(1) series x and a noise series y, as a
    # transfer-entropy job lays them out
    e = r.standard_normal(n + 1)
    xs = np.empty(n + 1)
    xs[0] = e[0]
    for t in range(1, n + 1):
        xs[t] = 0.8 * xs[t - 1] + 0.6 * e[t]
    x = np.stack([xs[:n], xs[1:], r.standard_normal(n)]).astype(np.float32)
    h = [0.25, 0.25, 0.3]
    ref = orc.density_fp64_window(x, n, [0, 1, 2], h)
    got, st = {}, {}
    try:
        for s_ in (0, sub):
            ctx.set_tuning(1, 32, mr, 64, s_)
            got[s_] = combined_densities(x, n, [0, 1, 2], h, ctx=ctx)
            st[s_] = ctx.last_stats()
            assert rel_err(got[s_], ref) < RTOL, s_
    finally:
        ctx.set_tuning()
    ev = {k: v["pairs_by_d"][2] for k, v in st.items()}
    assert 0 < ev[sub] < 0.9 * ev[0], ev


def test_sub_tiles_3d_batch_packed_and_sparse(ctx, orc, monkeypatch):
    """A batch of 1-D, 2-D and 3-D jobs at a tuning with sub-tiles: the 3-D
    jobs run as sub-tiles, the 2-D ones as whole tiles (the kernel variant
    with 3-D sub-tiles), the 1-D ones as usual.  Heavy-tailed 3-D data give
    short tiles, packed (ENT_TILE_PACK) into the sub-tile path with the
    offsets of the pack, and tiles wider than 32 bandwidths (raw-coordinate
    pair terms).  All match FP64, with packing on and off."""
    from entropical_amd import DeviceArray, batch_densities_dev

    r = np.random.default_rng(167)
    n = 20000
    x = np.concatenate([r.standard_normal((3, n)), r.standard_t(1.5, size=(3, n)) * 2.0]).astype(np.float32)
    jobs = [{"n_dim": 1, "n": n, "n_rows": n, "off": [0], "h": [0.1], "out_off": 0},
            {"n_dim": 2, "n": n, "n_rows": n, "off": [0, n], "h": [0.2, 0.25], "out_off": n},
            {"n_dim": 3, "n": n, "n_rows": n, "off": [0, n, 2 * n], "h": [0.3, 0.3, 0.35], "out_off": 2 * n},
            {"n_dim": 3, "n": n, "n_rows": n, "off": [3 * n, 4 * n, 5 * n], "h": [0.05, 0.4, 0.4],
             "out_off": 3 * n}]
    refs = [orc.density_fp64_window(x, n, [o // n for o in j["off"]], j["h"]) for j in jobs]
    dx = DeviceArray(ctx, x)
    out = ctx.alloc(8 * len(jobs) * n)
    try:
        for sub in (16, 32):
            ctx.set_tuning(1, 32, 1, 64, sub)
            for pack in ("0", "1"):
                monkeypatch.setenv("ENT_TILE_PACK", pack)
                batch_densities_dev(ctx, dx.ptr, jobs, out)
                got = np.empty(len(jobs) * n)
                ctx.d2h(got, out)
                for k, ref in enumerate(refs):
                    assert rel_err(got[k * n:(k + 1) * n], ref) < RTOL, (sub, pack, k)
            assert ctx.last_stats()["raw_tiles"] > 0
    finally:
        ctx.free(out)
        dx.free()
        ctx.set_tuning()


def test_moment_tables_over_cap_fall_back(ctx, orc, monkeypatch):
    """Moment tables that do not fit (here: a 1-byte cap) leave every pair to
    the pair loop: same densities within tolerance, no moment pairs, and the
    next run without the cap uses moments again."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(47)
    n = 60000
    a = r.standard_normal(n)
    x = np.stack([a, 0.5 * a + r.standard_normal(n)]).astype(np.float32)
    ref = orc.density_fp64(x, n, [0, 1], [0.2, 0.2])
    monkeypatch.setenv("ENT_MOMENT_TABLE_MAX", "1")
    capped = combined_densities(x, n, [0, 1], [0.2, 0.2], ctx=ctx)
    st0 = ctx.last_stats()
    monkeypatch.delenv("ENT_MOMENT_TABLE_MAX")
    again = combined_densities(x, n, [0, 1], [0.2, 0.2], ctx=ctx)
    st1 = ctx.last_stats()
    assert rel_err(capped, ref) <= RTOL
    assert rel_err(again, ref) <= RTOL
    assert st0["moment_pairs"] == 0
    assert st1["moment_pairs"] > 0
    assert st0["pairs"] == st1["pairs"]


@pytest.mark.parametrize("tuning", [(1, 32, 2, 64), (2, 16, 1, 16), (1, 32, 2, 64, 16), (1, 32, 1, 64, 32)])
def test_jobs_without_tables(ctx, orc, tuning):
    """Heavy-tailed data with narrow bandwidths have more rows of boxes than
    the table bound (8 entries per point + 4096): such jobs get neither a
    cell-start nor a row-start table, and every candidate range is a binary
    search over the job's keys.  2-D and 3-D jobs, whole tiles and
    sub-tiles, within 1e-5 of FP64."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(173)
    n = 3000
    x = (r.standard_t(1.5, size=(3, n)) * 3.0).astype(np.float32)
    m0, mr = tuning[1], tuning[2]
    try:
        ctx.set_tuning(*tuning)
        for cols, h in (([0, 1], [0.05, 0.01]), ([0, 1, 2], [0.05, 0.02, 0.02])):
            # the engine's box counts (ceil(range / (h / m)) per dim): the
            # rows of boxes exceed the table bound, so the job has no table
            nb = [int(np.ceil(np.float32(x[c].max() - x[c].min()) / np.float32(np.float32(hh) / (m0 if e == 0 else mr))))
                  for e, (c, hh) in enumerate(zip(cols, h))]
            assert int(np.prod(nb[1:])) + 1 > 8 * n + 4096, nb
            got = combined_densities(x, n, cols, h, ctx=ctx)
            assert rel_err(got, orc.density_fp64_window(x, n, cols, h)) < RTOL, (cols, h)
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("tuning", [(2, 32, 1, 64), (1, 32, 2, 64)])
def test_wide_sparse_tiles(ctx, orc, tuning):
    """Tiles in sparse stretches span up to a 64-bandwidth dim-0 group.  The
    local coordinates of such tiles must keep the pair terms exact enough:
    a narrow dim-0 bandwidth on normal data (whose tails give wide tiles;
    1.6e-5 with the tile's first query as origin) and heavy-tailed data
    (tiles wider than 32 bandwidths, raw-coordinate pair terms)."""
    from entropical_amd import combined_densities

    r = np.random.default_rng(109)
    ctx.set_tuning(*tuning)
    try:
        n = 20000
        x = r.standard_normal((2, n)).astype(np.float32)
        for h in ([0.0432, 0.15972], [0.03, 0.1]):
            got = combined_densities(x, n, [0, 1], h, ctx=ctx)
            assert rel_err(got, orc.density_fp64_window(x, n, [0, 1], h)) < RTOL, h
        y = (r.standard_t(1.5, size=(2, n)) * 3.0).astype(np.float32)
        for cols, h in (([0], [0.05]), ([0, 1], [0.05, 0.5])):
            got = combined_densities(y, n, cols, h, ctx=ctx)
            assert rel_err(got, orc.density_fp64_window(y, n, cols, h)) < RTOL, (cols, h)
            # the heavy tails give tiles wider than 32 bandwidths: the
            # raw-coordinate pair terms ran
            assert ctx.last_stats()["raw_tiles"] > 0, (cols, h)
        # dim-0 groups of 16 bandwidths bound every tile's span below 32
        # bandwidths: no tile takes the raw-coordinate path
        ctx.set_tuning(tuning[0], tuning[1], tuning[2], 16)
        for cols, h in (([0], [0.05]), ([0, 1], [0.05, 0.5])):
            got = combined_densities(y, n, cols, h, ctx=ctx)
            assert rel_err(got, orc.density_fp64_window(y, n, cols, h)) < RTOL, (cols, h)
            assert ctx.last_stats()["raw_tiles"] == 0, (cols, h)
    finally:
        ctx.set_tuning()


def test_key_width_change_replays_right_graph(ctx, orc, monkeypatch):
    """A shared-sort batch whose cell keys go from 64-bit (more than 2^32
    cells) to 32-bit and back at unchanged radix pass counts, on one context
    with the same coordinate and density buffers: each run's captured
    launch sequence must be the one of its own key width.  Jobs this sparse
    have a row-start table instead of a cell-start table, so the density
    kernel finds candidate ranges by binary search over a row's keys, which a
    key of the other width breaks."""
    from entropical_amd import DeviceArray, batch_densities_dev

    r = np.random.default_rng(131)
    n_base, n_cols = 1500, 4
    base = r.uniform(0.0, 1.0, size=(n_cols, n_base))
    # each point has a partner closer than every bandwidth used below
    x = np.concatenate([base, base + np.array([2e-5, 5e-6, 5e-6, 5e-6])[:, None]], axis=1)
    x = np.clip(x, 0.0, 1.0)
    x[:, 0], x[:, 1] = 0.0, 1.0  # range exactly 1 in every column
    x = x.astype(np.float32)
    n = x.shape[1]
    # m0 = 8, mr = 1: dim-0 boxes = 8 / h0, row boxes = 1 / h1.
    # wide:   2^17 x 2^16 cells (34 bits), passes 3 (dim 0) + 2 (rows)
    # narrow: 81920 x 32768 cells (32 bits), passes 3 + 2
    wide_h = (2.0 ** -14, 2.0 ** -16)
    narrow_h = (8.0 / 81920.0, 2.0 ** -15)
    monkeypatch.setenv("ENT_SORT", "shared")
    ctx.set_tuning(2, 8, 1, 4)
    dx = DeviceArray(ctx, x)
    out = ctx.alloc(8 * 3 * n)
    try:
        for h0, h1 in (wide_h, narrow_h, wide_h, narrow_h):
            jobs = [{"n_dim": 2, "n": n, "n_rows": n, "off": [0, c * n], "h": [h0, h1], "out_off": (c - 1) * n}
                    for c in (1, 2, 3)]
            batch_densities_dev(ctx, dx.ptr, jobs, out)
            got = np.empty(3 * n)
            ctx.d2h(got, out)
            for c in (1, 2, 3):
                ref = orc.density_fp64(x, n, [0, c], [h0, h1])
                assert rel_err(got[(c - 1) * n:c * n], ref) < RTOL, (h0, h1, c)
    finally:
        ctx.free(out)
        dx.free()
        ctx.set_tuning()


# ---------------------------------------------------------- epa convolution
def test_epa_convolution_matches_fp64(ctx, orc):
    from entropical_amd import epa_convolution

    r = np.random.default_rng(3)
    for n, h in ((1, 0.3), (2, 0.3), (5000, 0.05), (20000, 0.2)):
        x = np.sort(r.standard_normal(n).astype(np.float32))
        got = epa_convolution(x, h, ctx=ctx)
        ref = orc.epa_conv_fp64(x, h)
        assert got == pytest.approx(ref, rel=RTOL)
