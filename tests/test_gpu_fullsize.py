"""The benchmark workloads at their full sizes: the batched results against
the FP64 oracle on selected entries (the oracle's windowed O(N*W) density
keeps each check to seconds) and size-independent properties of the whole
result (symmetry, zero diagonal, finiteness, bit-identical repeats).
Tolerance 1e-4 relative per result, as for the small cases."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.fixture(scope="module")
def threads(orc):
    orc.set_threads(16)
    return 16


def test_mi_matrix_64x200000(ctx, orc, threads):
    from entropical_amd import DeviceArray, datasets, mi_matrix

    n_cols, n = 64, 200_000
    X = datasets.mixed_columns(n_cols, n)
    h = [0.15] * n_cols
    dx = DeviceArray(ctx, X)
    try:
        mi = mi_matrix(ctx, dx, h)
        again = mi_matrix(ctx, dx, h)
    finally:
        dx.free()
    np.testing.assert_array_equal(mi, again)
    assert np.all(np.isfinite(mi))
    np.testing.assert_array_equal(mi, mi.T)
    assert np.all(np.diag(mi) == 0)
    p = {c: orc.density_fp64_window(X, n, [c], [h[c]]) for c in (0, 1, 17, 42)}
    for i, j in ((0, 1), (17, 42)):
        pij = orc.density_fp64_window(X, n, [i, j], [h[i], h[j]])
        ref = orc.mi_fp64(pij, p[i], p[j], False)
        assert abs(mi[i, j] - ref) <= RTOL * abs(ref), (i, j, mi[i, j], ref)


@pytest.mark.parametrize("tuning", [(1, 32, 2, 64, 0), (1, 32, 3, 64, 16)])
def test_mi_densities_64x200000(ctx, orc, threads, tuning):
    """The MI batch at 64 x 200k with whole tiles and with 2-D sub-tiles of 16
    (cell and box moments active): per-point densities of three 2-D jobs and
    two 1-D jobs within 1e-5 of FP64 (mi.cpp:37-52 calls;
    densities_cuda.cu:46-105)."""
    from _jobsets import mi_jobs

    from entropical_amd import DeviceArray, batch_densities_dev, datasets

    n_cols, n = 64, 200_000
    X = datasets.mixed_columns(n_cols, n)
    h = [0.15] * n_cols
    jobs, slot, total = mi_jobs(n, n_cols, h)
    dx = DeviceArray(ctx, X)
    out = ctx.alloc(8 * total)
    keys = ((0, 1), (17, 42), (5, 63), (0,), (42,))
    got = {}
    try:
        ctx.set_tuning(*tuning)
        batch_densities_dev(ctx, dx.ptr, jobs, out)
        st = ctx.last_stats()
        for k in keys:
            g = np.empty(n)
            ctx.d2h(g, out + 8 * slot[k])
            got[k] = g
    finally:
        ctx.set_tuning()
        ctx.free(out)
        dx.free()
    # both moment paths ran in this batch
    assert st["moment_pairs_by_d"][0] > 0 and st["moment_pairs_by_d"][1] > 0
    for k in keys:
        ref = orc.density_fp64_window(X, n, list(k), [h[c] for c in k])
        err = rel_err(got[k], ref)
        assert err < 1e-5, (k, err)


def test_entropy1d_100000(ctx, orc, threads):
    """The 1-D entropy workload: N=100,000 draws of the three-component
    Gaussian mixture, h=0.1.  Every density within 1e-5 of FP64 (dens.cpp:
    20-46 through combined_densities), H = -mean ln p within 1e-4 of the
    oracle's FP64 value, at two tile shapes; and H close to the mixture's
    own differential entropy (a statistical check of the estimator)."""
    from entropical_amd import batch_entropy, combined_densities, datasets

    n, h = 100_000, 0.1
    x = datasets.gmm_1d(n)
    ref = orc.density_fp64_window(x, n, [0], [h])
    ref_H = orc.entropy(ref)
    ptr = ctx.alloc(x.nbytes)
    try:
        ctx.h2d(ptr, x)
        for tun in ((2, 8, 1, 4), (1, 32, 2, 64)):
            ctx.set_tuning(*tun)
            got = combined_densities(x, n, [0], [h], ctx=ctx)
            err = rel_err(got, ref)
            assert err < 1e-5, (tun, err)
            H = float(batch_entropy(ctx, ptr, [{"n_dim": 1, "n": n, "n_rows": n, "off": [0], "h": [h]}])[0])
            assert ctx.last_stats()["moment_pairs_by_d"][0] > 0
            assert abs(H - ref_H) <= RTOL * abs(ref_H), (tun, H, ref_H)
    finally:
        ctx.set_tuning()
        ctx.free(ptr)
    g = np.linspace(-6.0, 6.0, 240001)
    pdf = datasets.gmm_1d_pdf(g)
    H_true = float(-np.sum(pdf * np.log(np.maximum(pdf, 1e-300))) * (g[1] - g[0]))
    assert abs(ref_H - H_true) < 0.02, (ref_H, H_true)


@pytest.mark.parametrize("tuning", [(2, 16, 2, 16, 0), (1, 32, 1, 64, 16)])
def test_te_matrix_32x100000(ctx, orc, threads, tuning):
    """The 4-lag TE set at 32 x 100k, with whole tiles and with 3-D
    sub-tiles of 16 (the autotuned shape).  The same batch run through
    ent_batch_densities_dev gives a driven pair's 3-D densities
    p(x_t, x_t+tau, y_t) at every lag (and the reverse direction's at the
    largest lag) within 1e-5 of FP64 per point; the TE matrix entries of
    that pair match the oracle at every lag (transs.cpp:75-113)."""
    from _jobsets import te_jobs

    from entropical_amd import DeviceArray, batch_densities_dev, datasets, te_matrix

    X, drv, _ = datasets.coupled_ar1(32, 100_000)
    n = X.shape[1]
    h = [0.2] * 32
    taus = [1, 2, 5, 10]
    dx = DeviceArray(ctx, X)
    jobs, slot, total = te_jobs(n, 32, h, taus)
    out = ctx.alloc(8 * total)
    i = 5
    j = int(drv[i])
    a, b = min(i, j), max(i, j)
    dev = {}
    try:
        ctx.set_tuning(*tuning)
        te = te_matrix(ctx, dx, h, taus)
        batch_densities_dev(ctx, dx.ptr, jobs, out)
        for t, tau in enumerate(taus):
            g = np.empty(n - tau)
            ctx.d2h(g, out + 8 * slot[("itau", t, a, b)])
            dev[("itau", t)] = g
        g = np.empty(n - taus[-1])
        ctx.d2h(g, out + 8 * slot[("jtau", len(taus) - 1, a, b)])
        dev["jtau"] = g
    finally:
        ctx.set_tuning()
        ctx.free(out)
        dx.free()
    assert te.shape == (4, 32, 32) and np.all(np.isfinite(te))
    for t in range(4):
        assert np.all(np.diag(te[t]) == 0)
    f = lambda v: float(np.float32(v))
    pa = orc.density_fp64_window(X, n, [a], [h[a]])
    pb = orc.density_fp64_window(X, n, [b], [h[b]])
    pab = orc.density_fp64_window(X, n, [a, b], [h[a], h[b]])
    for t, tau in enumerate(taus):
        m = n - tau
        itau = orc.density_fp64_window(X, n, [a, a, b], [h[a], h[a], h[b]], [0, tau, 0])
        err = rel_err(dev[("itau", t)], itau)
        assert err < 1e-5, (tau, err)
        patau = orc.density_fp64_window(X, n, [a, a], [h[a], h[a]], [0, tau])
        ref_ab = orc.te_fp64(itau, pa[:m], pab[:m], patau, f(h[a]) * f(h[a]) * f(h[b]))
        assert abs(te[t, a, b] - ref_ab) <= RTOL * abs(ref_ab), (tau, a, b, te[t, a, b], ref_ab)
    tau = taus[-1]
    m = n - tau
    jtau = orc.density_fp64_window(X, n, [a, b, b], [h[a], h[b], h[b]], [0, 0, tau])
    err = rel_err(dev["jtau"], jtau)
    assert err < 1e-5, err
    pbtau = orc.density_fp64_window(X, n, [b, b], [h[b], h[b]], [0, tau])
    ref_ba = orc.te_fp64(jtau, pb[:m], pab[:m], pbtau, f(h[a]) * f(h[b]) * f(h[a]))
    assert abs(te[-1, b, a] - ref_ba) <= RTOL * abs(ref_ba), (b, a, te[-1, b, a], ref_ba)


def test_uneven_batch_500(ctx, orc, threads):
    from entropical_amd import batch_entropy, datasets

    buf, specs = datasets.uneven_batch()
    ptr = ctx.alloc(buf.nbytes)
    try:
        ctx.h2d(ptr, buf)
        H = batch_entropy(ctx, ptr, specs)
        again = batch_entropy(ctx, ptr, specs)
    finally:
        ctx.free(ptr)
    np.testing.assert_array_equal(H, again)
    assert np.all(np.isfinite(H))
    # the longest trajectory of each dimension
    for d in (1, 2, 3):
        k = max((s for s in range(len(specs)) if specs[s]["n_dim"] == d), key=lambda s: specs[s]["n"])
        s = specs[k]
        pts = np.stack([buf[s["off"][e]: s["off"][e] + s["n"]] for e in range(d)])
        p = orc.density_fp64_window(pts, s["n"], list(range(d)), s["h"][:d])
        ref = orc.entropy(p)
        assert abs(H[k] - ref) <= RTOL * abs(ref), (k, d, H[k], ref)
