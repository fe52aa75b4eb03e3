"""GPU parity of the batched drivers (MI matrix, transfer entropy, entropy)
against the FP64 oracle: oracle densities + the oracle's FP64 restatement of
each reference formula.  Tolerance 1e-4 relative per result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def coupled(seed, c, n):
    r = np.random.default_rng(seed)
    x = r.standard_normal((c, n))
    for k in range(1, c):
        x[k] += 0.6 * x[k - 1]
    return x.astype(np.float32)


@pytest.mark.parametrize("method", ["cell", "brute"])
@pytest.mark.parametrize("bits", [False, True])
def test_mi_matrix(ctx, orc, method, bits):
    from entropical_amd import DeviceArray, mi_matrix

    c, n = 5, 4000
    x = coupled(1, c, n)
    h = [0.3, 0.25, 0.35, 0.2, 0.4]
    dx = DeviceArray(ctx, x)
    try:
        got = mi_matrix(ctx, dx, h, bits=bits, method=method)
    finally:
        dx.free()
    p = [orc.density_fp64(x, n, [i], [h[i]]) for i in range(c)]
    assert np.all(np.diag(got) == 0)
    for i in range(c):
        for j in range(i + 1, c):
            pij = orc.density_fp64(x, n, [i, j], [h[i], h[j]])
            ref = orc.mi_fp64(pij, p[i], p[j], bits)
            assert got[i, j] == got[j, i]
            assert abs(got[i, j] - ref) <= RTOL * abs(ref), (i, j, got[i, j], ref)


def test_graph_replays_match_eager(ctx):
    """The captured preprocessing + density graphs, replayed several times with
    eager work in between, give bit-identical results to eager launches."""
    from entropical_amd import DeviceArray, combined_densities, mi_matrix

    c, n = 6, 6000
    x = coupled(5, c, n)
    h = [0.3] * c
    dx = DeviceArray(ctx, x)
    try:
        ctx.set_graph(False)
        eager = mi_matrix(ctx, dx, h)
        ctx.set_graph(True)
        for rep in range(4):
            got = mi_matrix(ctx, dx, h)
            np.testing.assert_array_equal(got, eager)
            if rep == 1:
                assert ctx.last_stats()["density_ms"] > 0
            # unrelated eager work between replays (different plan shape)
            combined_densities(x, n, [0, 1, 2], [0.4] * 3, ctx=ctx)
    finally:
        ctx.set_graph(True)
        dx.free()


def test_graph_replay_on_changed_data(ctx):
    """A step replayed from the captured graphs on changed data of the same
    shape (new values written into the same device buffer) equals an eager
    run on that data bit for bit.  The second data set permutes each
    column's rows (same value range, so the same plan shape and a replay of
    the graphs captured for the first, with different MI); the third is
    rescaled (a new plan shape, captured anew)."""
    from entropical_amd import DeviceArray, mi_matrix

    c, n = 5, 8000
    h = [0.25] * c
    base = coupled(21, c, n)
    r = np.random.default_rng(22)
    perm = np.stack([base[k, r.permutation(n)] for k in range(c)])
    sets = [base, perm, (1.7 * coupled(23, c, n)).astype(np.float32)]
    dx = DeviceArray(ctx, base)
    try:
        prev = None
        for k, x in enumerate(sets):
            ctx.h2d(dx.ptr, x)
            ctx.set_graph(False)
            eager = mi_matrix(ctx, dx, h)
            ctx.set_graph(True)
            first = mi_matrix(ctx, dx, h)
            replays_first = ctx.last_stats()["graph_replays"]
            again = mi_matrix(ctx, dx, h)
            assert ctx.last_stats()["graph_replays"] >= 1
            np.testing.assert_array_equal(first, eager)
            np.testing.assert_array_equal(again, eager)
            if k == 1:
                # same shape as the first set: its graphs replayed on new values
                assert replays_first >= 1
                assert not np.array_equal(eager, prev)
            prev = eager
    finally:
        ctx.set_graph(True)
        dx.free()


def test_te_matrix(ctx, orc):
    from entropical_amd import DeviceArray, te_matrix

    c, n = 3, 3000
    x = coupled(2, c, n)
    h = [0.35, 0.3, 0.4]
    taus = [1, 4]
    dx = DeviceArray(ctx, x)
    try:
        got = te_matrix(ctx, dx, h, taus)
    finally:
        dx.free()
    f = lambda v: float(np.float32(v))
    p = [orc.density_fp64(x, n, [i], [h[i]]) for i in range(c)]
    for t, tau in enumerate(taus):
        ptau = [orc.density_fp64(x, n, [i, i], [h[i], h[i]], [0, tau]) for i in range(c)]
        for i in range(c):
            for j in range(i + 1, c):
                pij = orc.density_fp64(x, n, [i, j], [h[i], h[j]])
                itau = orc.density_fp64(x, n, [i, i, j], [h[i], h[i], h[j]], [0, tau, 0])
                jtau = orc.density_fp64(x, n, [i, j, j], [h[i], h[j], h[j]], [0, 0, tau])
                m = n - tau
                ref_ij = orc.te_fp64(itau, p[i][:m], pij[:m], ptau[i], f(h[i]) * f(h[i]) * f(h[j]))
                ref_ji = orc.te_fp64(jtau, p[j][:m], pij[:m], ptau[j], f(h[i]) * f(h[j]) * f(h[i]))
                assert abs(got[t, i, j] - ref_ij) <= RTOL * abs(ref_ij) + 1e-12, (tau, i, j)
                assert abs(got[t, j, i] - ref_ji) <= RTOL * abs(ref_ji) + 1e-12, (tau, j, i)
        assert np.all(np.diag(got[t]) == 0)


def test_batch_entropy_uneven_jobs(ctx, orc):
    from entropical_amd import batch_entropy

    r = np.random.default_rng(3)
    lens = [1500, 4000, 777, 2500]
    dims = [1, 2, 3, 2]
    hs = [0.1, 0.25, 0.4, 0.15]
    blocks, specs, off = [], [], 0
    for L, d, hh in zip(lens, dims, hs):
        blk = r.standard_t(4, size=(d, L)).astype(np.float32)
        blocks.append(blk)
        specs.append({"n_dim": d, "n": L, "n_rows": L, "off": [off + e * L for e in range(d)], "h": [hh] * d})
        off += d * L
    flat = np.concatenate([b.reshape(-1) for b in blocks])
    ptr = ctx.alloc(flat.nbytes)
    try:
        ctx.h2d(ptr, flat)
        got = batch_entropy(ctx, ptr, specs)
    finally:
        ctx.free(ptr)
    for k, (blk, d, hh) in enumerate(zip(blocks, dims, hs)):
        p = orc.density_fp64(blk, blk.shape[1], list(range(d)), [hh] * d)
        ref = orc.entropy(p)
        assert abs(got[k] - ref) <= RTOL * abs(ref), (k, got[k], ref)


def _shared_window_batch(c, n, lags):
    """MI-style 1-D/2-D jobs plus TE-style lagged 3-D jobs over c columns of
    n rows (column k at offset k*n): many jobs share a dim-0 window."""
    specs, out = [], 0
    for i in range(c):
        specs.append({"n_dim": 1, "n": n, "n_rows": n, "off": [i * n], "h": [0.2 + 0.05 * i]})
        for j in range(i + 1, c):
            specs.append({"n_dim": 2, "n": n, "n_rows": n, "off": [i * n, j * n], "h": [0.2 + 0.05 * i, 0.3]})
    for tau in lags:
        m = n - tau
        for i in range(c):
            for j in range(c):
                if i != j:
                    specs.append({"n_dim": 3, "n": m, "n_rows": n, "off": [i * n, i * n + tau, j * n],
                                  "h": [0.2 + 0.05 * i, 0.25, 0.35]})
    for s in specs:
        s["out_off"] = out
        out += s["n"]
    return specs, out


def _batch(ctx, ptr, specs, total):
    from entropical_amd import batch_densities_dev

    dptr = ctx.alloc(8 * total)
    try:
        batch_densities_dev(ctx, ptr, specs, dptr)
        got = np.empty(total, dtype=np.float64)
        ctx.d2h(got, dptr)
    finally:
        ctx.free(dptr)
    return got


@pytest.mark.parametrize("sort_mode", ["full", "shared"])
@pytest.mark.parametrize("scatter", ["direct", "staged"])
def test_sort_variants_bit_identical(ctx, orc, monkeypatch, sort_mode, scatter):
    """One sort per job by its full cell key, or shared dim-0 sorts followed
    by passes over the remaining dims' key, with either scatter kernel, all
    give the same permutation: densities are bit-identical to the default
    run's, and that run matches the FP64 oracle."""
    c, n = 4, 12000
    x = coupled(11, c, n)
    specs, total = _shared_window_batch(c, n, [1, 5])
    ptr = ctx.alloc(x.nbytes)
    try:
        ctx.h2d(ptr, x)
        ref_bits = _batch(ctx, ptr, specs, total)
        # auto picks the shared sort here: one base per (column, window length)
        assert ctx.last_stats()["sort_bases"] == c * 3
        monkeypatch.setenv("ENT_SORT", sort_mode)
        monkeypatch.setenv("ENT_SCATTER", scatter)
        got = _batch(ctx, ptr, specs, total)
        st = ctx.last_stats()
    finally:
        ctx.free(ptr)
    assert (st["sort_bases"] > 0) == (sort_mode == "shared")
    np.testing.assert_array_equal(got, ref_bits)
    for k in (0, 1, len(specs) - 1):
        s = specs[k]
        cols = [o // n for o in s["off"]]
        taus = [o % n for o in s["off"]]
        p = orc.density_fp64(x, n, cols, s["h"], taus if s["n_dim"] == 3 else None)
        g = got[s["out_off"]: s["out_off"] + s["n"]]
        assert np.max(np.abs(g - p) / p) < 1e-5, k


# ------------------------------------------------- autoent / kldiv / negs
def _distinct_columns(seed, n):
    r = np.random.default_rng(seed)
    x = np.stack([
        r.standard_normal(n),
        np.where(r.random(n) < 0.4, -1.5, 1.0) + 0.5 * r.standard_normal(n),
        0.7 * r.standard_t(5, n),
        np.cumsum(r.standard_normal(n)) * 0.02 + r.standard_normal(n),  # autocorrelated
    ])
    return x.astype(np.float32)


@pytest.mark.parametrize("bits", [False, True])
def test_autoent(ctx, orc, bits):
    from entropical_amd import DeviceArray, autoent

    n, tau = 5000, 3
    x = _distinct_columns(21, n)
    h = [0.3, 0.25, 0.35, 0.3]
    dx = DeviceArray(ctx, x)
    try:
        got = autoent(ctx, dx, h, tau, bits=bits)
    finally:
        dx.free()
    for i in range(4):
        p1 = orc.density_fp64(x, n, [i], [h[i]], [tau])
        p2 = orc.density_fp64(x, n, [i, i], [h[i], h[i]], [0, tau])
        ref = orc.autoent_ref(p1, p2, h[i], bits)
        assert abs(got[i] - ref) <= RTOL * abs(ref), (i, got[i], ref)


@pytest.mark.parametrize("bits", [False, True])
def test_kldiv_and_negs(ctx, orc, bits):
    from entropical_amd import DeviceArray, kldiv, negs

    n = 6000
    x = _distinct_columns(22, n)
    h = [0.3, 0.25, 0.35, 0.3]
    dx = DeviceArray(ctx, x)
    try:
        kl = kldiv(ctx, dx, h, bits=bits)
        ng = negs(ctx, dx, h, bits=bits)
    finally:
        dx.free()
    p = [orc.sum1_normalized(orc.density_fp64(x, n, [i], [h[i]])) for i in range(4)]
    assert np.all(np.diag(kl) == 0)
    for i in range(4):
        ref = orc.negs_ref(p[i], bits)
        assert abs(ng[i] - ref) <= RTOL * abs(ref), (i, ng[i], ref)
        for j in range(i + 1, 4):
            ij, ji = orc.kldiv_ref(p[i], p[j], bits)
            assert abs(kl[i, j] - ij) <= RTOL * abs(ij), (i, j, kl[i, j], ij)
            assert abs(kl[j, i] - ji) <= RTOL * abs(ji), (j, i, kl[j, i], ji)
