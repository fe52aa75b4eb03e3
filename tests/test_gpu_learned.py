"""Learned bandwidth on the GPU: MLP kernels against the FP64 oracle, LOO
labels against the CPU oracle, and the gain of the trained network over the
rule of thumb (hestimate.cpp:19-30) on held-out synthetic samples."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_fp64_oracle(ctx):
    from entropical_amd.learned_h import BandwidthModel, HIDDEN, HIDDEN_LAYERS, N_FEATURES
    from oracle import mlp_oracle as M

    r = np.random.default_rng(0)
    model = BandwidthModel(ctx, seed=3)
    p = model.get_params()
    X = r.normal(size=(96, N_FEATURES)).astype(np.float32)
    t = r.normal(size=96).astype(np.float32)
    g, loss = model.gradient_raw(X, t)
    dims = [N_FEATURES] + [HIDDEN] * HIDDEN_LAYERS + [1]
    ref_loss, ref_g = M.loss_and_grad(p.astype(np.float64), dims, X.astype(np.float64), t, bf16=True)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    # FP32 accumulation vs FP64: compare relative to the gradient scale
    err = np.abs(g - ref_g) / (np.abs(ref_g) + 1e-3 * np.abs(ref_g).max())
    assert np.max(err) < 1e-3
    assert np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g) < 1e-5
    model.close()


def test_mlp_forward_matches_oracle(ctx):
    from entropical_amd.learned_h import BandwidthModel, HIDDEN, HIDDEN_LAYERS, N_FEATURES
    from oracle import mlp_oracle as M

    r = np.random.default_rng(1)
    model = BandwidthModel(ctx, seed=4)
    X = r.normal(size=(300, N_FEATURES)).astype(np.float32)
    y = model.predict_raw(X)
    dims = [N_FEATURES] + [HIDDEN] * HIDDEN_LAYERS + [1]
    ref = M.forward(model.get_params().astype(np.float64), dims, X.astype(np.float64))[-1][:, 0]
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-5)
    model.close()


def test_mlp_training_reduces_loss(ctx):
    from entropical_amd.learned_h import BandwidthModel, N_FEATURES

    r = np.random.default_rng(2)
    X = r.normal(size=(2048, N_FEATURES)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    model = BandwidthModel(ctx, seed=5)
    losses = model.fit(X, t, epochs=40, batch=128, lr=2e-3)
    assert losses[-1] < 0.2 * losses[0]
    Xv = r.normal(size=(512, N_FEATURES)).astype(np.float32)
    tv = np.sin(Xv[:, 0]) + 0.5 * Xv[:, 1] * Xv[:, 2]
    pv = model.predict_raw(model._norm(Xv))
    assert np.mean((pv - tv) ** 2) < 0.3 * np.var(tv)
    model.close()


def test_loo_labels_match_oracle(ctx, orc):
    from entropical_amd.learned_h import FLOOR_REL, loo_grid, thumb

    r = np.random.default_rng(6)
    ds = [r.standard_normal((d, n)).astype(np.float32) * [[1.0], [2.0], [0.5]][:d] for d, n in ((1, 700), (2, 500), (3, 400))]
    grid = np.array([0.25, 0.7, 1.0, 2.0])
    ll = loo_grid(ctx, ds, r_grid=grid)
    for k, x in enumerate(ds):
        d, n = x.shape
        th = thumb(x)
        for c, rr in enumerate(grid):
            h = (rr * th).astype(np.float32)  # the float32 bandwidths the jobs carry
            k0 = np.prod(0.75 / h.astype(np.float64))
            ref = orc.loo_loglik(x.T.reshape(-1), d, h, [1.0], FLOOR_REL * k0 / (n - 1))[0]
            assert ll[k, c] == pytest.approx(ref, rel=1e-5, abs=1e-5), (k, c)


def test_trained_network_beats_rule_of_thumb(ctx):
    """On held-out synthetic samples (a seed unseen in training) the network's
    bandwidth reaches a higher leave-one-out likelihood than the rule of thumb
    it corrects, and lands closer to the LOO-optimal bandwidth."""
    from entropical_amd import learned_h as L

    model = L.BandwidthModel.load(ctx)
    rng = np.random.default_rng(123456789)
    ds = [L.synth_dataset(rng, 1000, 12000) for _ in range(48)]
    ll = L.loo_grid(ctx, ds)
    F = np.stack([L.features(x) for x in ds])
    pred = model.predict_log2r(F)
    lr = np.log2(L.R_GRID)
    best = np.array([L.label_from_grid(row) for row in ll])
    ll_mlp = np.array([np.interp(p, lr, row) for p, row in zip(pred, ll)])
    ll_thumb = np.array([np.interp(0.0, lr, row) for row in ll])
    n = np.array([x.shape[1] for x in ds])
    # mean per-point LOO log-likelihood gain
    gain = np.mean((ll_mlp - ll_thumb) / n)
    assert gain > 0.0
    assert np.median(np.abs(pred - best)) < np.median(np.abs(best))
    model.close()


def test_column_stats_match_host(ctx):
    """ent_column_stats_dev (the learned bandwidth's per-column statistics on
    the device) against the host restatement column_stats: the 18 order
    statistics bit for bit (radix select), the tail count and the z histogram
    exactly, mean and variance to 1e-12, the z sums to 1e-9; the assembled
    features agree to 1e-6.  Columns: normal, Student-t(2) tails, a bimodal
    mixture, heavy ties, an offset tiny-scale column, mixed signs with zeros,
    and a constant column; lengths 1000 and 1001."""
    from entropical_amd import DeviceArray
    from entropical_amd import learned_h as L

    r = np.random.default_rng(71)
    for n in (1000, 1001, 37_123):
        cols = [r.standard_normal(n), r.standard_t(2, n), np.where(r.random(n) < 0.3, -2.0, 1.5) + 0.2 * r.standard_normal(n),
                np.round(r.standard_normal(n) * 3.0) / 3.0, 1e3 + 1e-3 * r.standard_normal(n),
                np.where(r.random(n) < 0.2, 0.0, r.standard_normal(n)) * np.where(r.random(n) < 0.5, -1.0, 1.0),
                np.full(n, 2.5)]
        X = np.ascontiguousarray(np.stack(cols).astype(np.float32))
        X[5, :5] = -0.0
        dx = DeviceArray(ctx, X)
        try:
            dev = L.column_stats_dev(ctx, dx.ptr, n, X.shape[0])
            F_dev, _ = L.column_features_dev(ctx, dx.ptr, n, X.shape[0])
        finally:
            dx.free()
        for c in range(X.shape[0]):
            host = L.column_stats(X[c])
            k = 2 + L.N_RANKS
            np.testing.assert_array_equal(dev[c, 2:k], host[2:k], err_msg=f"order statistics n={n} col={c}")
            np.testing.assert_array_equal(dev[c, k + 3:], host[k + 3:], err_msg=f"counts n={n} col={c}")
            np.testing.assert_allclose(dev[c, :2], host[:2], rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(dev[c, k:k + 3], host[k:k + 3], rtol=1e-9, atol=1e-9 * n)
            np.testing.assert_allclose(F_dev[c], L.features(X[c:c + 1]), atol=1e-6, err_msg=f"n={n} col={c}")


def test_network_on_mi_columns_against_hestimate(ctx):
    """The MI workload's 64 x 200,000 columns: the network's bandwidths
    (features' statistics and the network on the device) give a leave-one-out
    log-likelihood closer to the best of the 32-multiplier grid than both
    hestimate bandwidths (rule of thumb and AMISE, computed on the GPU), and
    the learned pipeline takes less time than AMISE."""
    import time

    from entropical_amd import DeviceArray, datasets, hestimate
    from entropical_amd import learned_h as L

    X = datasets.mixed_columns(64, 200_000)
    n_cols, n = X.shape
    model = L.BandwidthModel.load(ctx)
    dx = DeviceArray(ctx, X)
    try:
        L.column_bandwidths(ctx, model, dx)  # first call: module and buffer set-up
        t0 = time.perf_counter()
        h_net, used = L.column_bandwidths(ctx, model, dx)
        t_net = time.perf_counter() - t0
        assert used.all()
        h_th = hestimate(ctx, dx, "thumb")
        t0 = time.perf_counter()
        h_am = hestimate(ctx, dx, "amise")
        t_am = time.perf_counter() - t0
        best = L.loo_grid(ctx, [X[c: c + 1] for c in range(n_cols)]).max(axis=1) / n

        def loo(h):
            specs = [{"n_dim": 1, "n": n, "n_rows": n, "off": [c * n], "h": [float(h[c])]} for c in range(n_cols)]
            return L.batch_loo(ctx, dx.ptr, specs) / n

        gap = {k: float(np.mean(best - loo(h))) for k, h in (("net", h_net), ("thumb", h_th), ("amise", h_am))}
    finally:
        dx.free()
        model.close()
    assert gap["net"] < gap["thumb"] and gap["net"] < gap["amise"], gap
    assert t_net < t_am, (t_net, t_am)
