"""The benchmark workloads at their full sizes: the batched results against
the FP64 oracle on selected entries (the oracle's windowed O(N*W) density
keeps each check to seconds) and size-independent properties of the whole
result (symmetry, zero diagonal, finiteness, bit-identical repeats).
Tolerance 1e-4 relative per result, as for the small cases."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-4


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


def test_te_matrix_32x100000(ctx, orc, threads):
    from entropical_amd import DeviceArray, datasets, te_matrix

    X, drv, _ = datasets.coupled_ar1(32, 100_000)
    n = X.shape[1]
    h = [0.2] * 32
    taus = [1, 2, 5, 10]
    dx = DeviceArray(ctx, X)
    try:
        te = te_matrix(ctx, dx, h, taus)
    finally:
        dx.free()
    assert te.shape == (4, 32, 32) and np.all(np.isfinite(te))
    for t in range(4):
        assert np.all(np.diag(te[t]) == 0)
    # one driven pair (j drives i), checked at lag 1 in both directions
    i = 5
    j = int(drv[i])
    a, b = min(i, j), max(i, j)
    tau = 1
    f = lambda v: float(np.float32(v))
    m = n - tau
    pa = orc.density_fp64_window(X, n, [a], [h[a]])
    pb = orc.density_fp64_window(X, n, [b], [h[b]])
    pab = orc.density_fp64_window(X, n, [a, b], [h[a], h[b]])
    patau = orc.density_fp64_window(X, n, [a, a], [h[a], h[a]], [0, tau])
    pbtau = orc.density_fp64_window(X, n, [b, b], [h[b], h[b]], [0, tau])
    itau = orc.density_fp64_window(X, n, [a, a, b], [h[a], h[a], h[b]], [0, tau, 0])
    jtau = orc.density_fp64_window(X, n, [a, b, b], [h[a], h[b], h[b]], [0, 0, tau])
    ref_ab = orc.te_fp64(itau, pa[:m], pab[:m], patau, f(h[a]) * f(h[a]) * f(h[b]))
    ref_ba = orc.te_fp64(jtau, pb[:m], pab[:m], pbtau, f(h[a]) * f(h[b]) * f(h[a]))
    assert abs(te[0, a, b] - ref_ab) <= RTOL * abs(ref_ab), (a, b, te[0, a, b], ref_ab)
    assert abs(te[0, b, a] - ref_ba) <= RTOL * abs(ref_ba), (b, a, te[0, b, a], ref_ba)


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
