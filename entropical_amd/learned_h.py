"""Learned bandwidth selection.

Bandwidth family: per dimension e, ``h_e = r * thumb_e`` with the reference's
rule of thumb ``thumb_e = N^(-1/7) * sigma_e`` (Hestimate::Thumb::Kernel,
hestimate.cpp:19-30).  A small MLP (24 features -> 128 -> 128 -> 128 -> 1,
tanh, BF16 weights, FP32 accumulation; trained and evaluated by the HIP
kernels in mlp.hip) predicts ``log2 r``.  Labels are the leave-one-out
log-likelihood argmax over 32 candidate multipliers, computed by the GPU KDE
(``ent_batch_loo_dev``).  The rule of thumb (and AMISE through
``epa_convolution``) remain the fallback and the yardstick.

Training data: seeded synthetic samples with d in 1..3, N log-uniform in
[1000, 50000], 1-5 Gaussian or Student-t components with random
covariances.  The per-feature range of the training set is stored with the
weights; inputs whose shape features fall outside it get the AMISE bandwidth
(hestimate) instead of the network's.  The sample size is not part of that
check: the rule of thumb carries the N^(-1/7) scaling, and the network
extrapolates in log10 N beyond the training range (the MI workload's columns
have N = 200,000).
"""
from __future__ import annotations

import ctypes as C
import json
import time
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, load
from .kde import Context, _ptr, make_jobs

N_FEATURES = 24
HIDDEN = 128
HIDDEN_LAYERS = 3
R_GRID = 2.0 ** np.linspace(-3.0, 2.0, 32)  # candidate multipliers of the rule of thumb
FLOOR_REL = 1e-3                            # LOO floor: 1/1000 of one coincident neighbour
WEIGHTS = Path(__file__).resolve().parent / "learned_h_weights.npz"


# ----------------------------------------------------------------- data
N_MIN, N_MAX = 1000, 50_000


def synth_dataset(rng: np.random.Generator, n_min: int = N_MIN, n_max: int = N_MAX) -> np.ndarray:
    """One training sample: (d, N) float32 points from a random mixture."""
    d = int(rng.integers(1, 4))
    n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
    k = int(rng.integers(1, 6))
    w = rng.dirichlet(np.ones(k))
    comp = rng.choice(k, size=n, p=w)
    x = np.empty((n, d))
    for c in range(k):
        m = comp == c
        cnt = int(m.sum())
        if cnt == 0:
            continue
        mean = rng.normal(0.0, 2.0, d)
        q, _ = np.linalg.qr(rng.normal(size=(d, d)))
        L = q * rng.uniform(0.2, 1.5, d)
        z = rng.standard_normal((cnt, d))
        if rng.random() < 0.5:  # Student-t component
            nu = float(rng.integers(3, 11))
            z *= np.sqrt(nu / rng.chisquare(nu, size=(cnt, 1)))
        x[m] = mean + z @ L.T
    return np.ascontiguousarray(x.T.astype(np.float32))


def thumb(x: np.ndarray) -> np.ndarray:
    """Rule-of-thumb bandwidth per dimension, N^(-1/7) * population sd
    (hestimate.cpp:19-30, variance in FP64)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[1]
    return (n ** (-1.0 / 7.0)) * x.std(axis=1)


_QS = np.array([0.01, 0.05, 0.10, 0.25, 0.50, 0.75, 0.90, 0.95, 0.99])
N_RANKS = 2 * len(_QS)  # order statistics per dimension (ENT_COLSTATS_RANKS)
N_HIST = 32             # bins of the z histogram over [-4, 4]


def quantile_ranks(n: int):
    """0-based ranks of the order statistics the quantiles interpolate
    between, and the interpolation weights, as numpy's 'linear' quantile
    computes them: virtual index n q + (1 - q) - 1 = q (n - 1)."""
    vi = n * _QS + (1.0 + _QS * (1.0 - 1.0 - 1.0)) - 1.0
    lo = np.floor(vi)
    gamma = vi - lo
    lo = lo.astype(np.int64).clip(0, n - 1)
    hi = (lo + 1).clip(0, n - 1)
    return np.concatenate([lo, hi]).astype(np.uint32), gamma


def _lerp(a, b, t):
    """numpy's quantile interpolation: a + (b - a) t, or b - (b - a)(1 - t)
    from t = 1/2 on."""
    d = b - a
    return np.where(t >= 0.5, b - d * (1.0 - t), a + d * t)


def column_stats(x: np.ndarray) -> np.ndarray:
    """Sufficient statistics of one dimension's sample (1-D float32) for the
    features, host version of ent_column_stats_dev: mean, population
    variance, the order statistics at quantile_ranks, and with z the z-score:
    sum z^3, sum z^4, sum |z|, count |z| > 3, 32-bin histogram of z over
    [-4, 4].  Layout of one ENT_COLSTATS_WIDTH row."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    mu = x.mean()
    var = x.var()
    sd = np.sqrt(var) if var > 0 else 1.0
    z = (x - mu) / sd
    ranks, _ = quantile_ranks(n)
    os_ = np.partition(x, np.unique(ranks))[ranks]
    hcount, _ = np.histogram(z, bins=N_HIST, range=(-4, 4))
    return np.concatenate([[mu, var], os_, [(z ** 3).sum(), (z ** 4).sum(), np.abs(z).sum(),
                                            float(np.count_nonzero(np.abs(z) > 3.0))], hcount.astype(np.float64)])


def features_from_stats(st: np.ndarray, n: int, corr: Optional[np.ndarray] = None) -> np.ndarray:
    """The 24 features of a (d, N) sample from its dimensions' statistics
    (st: d rows of column_stats layout) and, for d > 1, the correlation
    matrix of its dimensions."""
    st = np.atleast_2d(np.asarray(st, dtype=np.float64))
    d = st.shape[0]
    mu, var = st[:, 0], st[:, 1]
    sd = np.sqrt(var)
    sd[var == 0] = 1.0
    _, gamma = quantile_ranks(n)
    nq = len(_QS)
    zlo = (st[:, 2:2 + nq] - mu[:, None]) / sd[:, None]
    zhi = (st[:, 2 + nq:2 + 2 * nq] - mu[:, None]) / sd[:, None]
    q = _lerp(zlo, zhi, gamma[None, :]).T  # (9, d)
    k = 2 + N_RANKS
    skew = st[:, k] / n
    kurt = st[:, k + 1] / n - 3.0
    iqr = q[5] - q[3]
    f = np.zeros(N_FEATURES)
    f[0] = np.log10(n)
    f[1] = d
    f[2:11] = q.mean(axis=1)
    f[11] = np.abs(skew).mean()
    f[12] = np.clip(kurt, -2.0, 20.0).mean() / 10.0
    f[13] = iqr.mean()
    f[14] = (q[8] - q[0]).mean()
    # bimodality coefficient; a constant dimension (fourth moment 0) counts 0
    m4 = kurt + 3.0
    f[15] = np.divide(skew ** 2 + 1.0, m4, out=np.zeros_like(m4), where=m4 > 0).mean()
    if d > 1:
        off = np.abs(corr[~np.eye(d, dtype=bool)])
        f[16] = off.mean()
        f[17] = off.max()
        f[18] = max(np.log(max(np.linalg.det(corr), 1e-12)), -10.0) / 10.0
    f[19] = np.minimum(1.0, iqr / 1.349).mean()
    f[20] = st[:, k + 3].sum() / (d * n) * 10.0
    f[21] = (st[:, k + 2] / n).mean()
    f[22] = np.log2(n ** (-1.0 / (d + 4)) / n ** (-1.0 / 7.0))
    modes = []
    for e in range(d):
        hcount = st[e, k + 4:k + 4 + N_HIST]
        hs = np.convolve(hcount, np.ones(3) / 3.0, mode="same")
        modes.append(np.sum((hs[1:-1] > hs[:-2]) & (hs[1:-1] >= hs[2:]) & (hs[1:-1] > 0.01 * hs.max())))
    f[23] = np.mean(modes) / 5.0
    return f.astype(np.float32)


def features(x: np.ndarray) -> np.ndarray:
    """24 sample statistics of a (d, N) sample (scale-free: per-dim z-scores):
    mean quantiles of the z-scores, skewness, kurtosis, spreads, tail mass,
    modality, correlation structure for d > 1, and N, d."""
    x = np.asarray(x, dtype=np.float64)
    d, n = x.shape
    st = np.stack([column_stats(x[e]) for e in range(d)])
    corr = np.corrcoef(x) if d > 1 else None
    return features_from_stats(st, n, corr)


def column_stats_dev(ctx: Context, x_dev: int, n_rows: int, n_cols: int) -> np.ndarray:
    """column_stats of every column of device-resident column-major X,
    computed on the device (ent_column_stats_dev): (n_cols, 56) float64."""
    ranks, _ = quantile_ranks(n_rows)
    ranks = np.ascontiguousarray(ranks, dtype=np.uint32)
    st = np.empty((n_cols, 2 + N_RANKS + 4 + N_HIST), np.float64)
    check(load().ent_column_stats_dev(ctx.handle, C.c_void_p(x_dev), n_rows, n_cols, _ptr(ranks), _ptr(st)))
    return st


def column_features_dev(ctx: Context, x_dev: int, n_rows: int, n_cols: int):
    """The features of every column (as a 1-D sample) of device-resident
    column-major X: the per-column statistics on the device, the 24 numbers
    per column assembled from them here.  Returns (features (n_cols, 24),
    statistics (n_cols, 56))."""
    st = column_stats_dev(ctx, x_dev, n_rows, n_cols)
    return np.stack([features_from_stats(st[c:c + 1], n_rows) for c in range(n_cols)]), st


# --------------------------------------------------------------- labels
def batch_loo(ctx: Context, x_dev_ptr: int, specs: Sequence[dict], floor_rel: float = FLOOR_REL) -> np.ndarray:
    """Leave-one-out log-likelihood (sum over points) of every job, GPU KDE."""
    arr = make_jobs(specs)
    res = np.empty(len(specs))
    check(load().ent_batch_loo_dev(ctx.handle, C.c_void_p(x_dev_ptr), arr, len(specs), float(floor_rel),
                                   _lib.METHOD_CELL, _ptr(res)))
    return res


def loo_grid(ctx: Context, datasets: Sequence[np.ndarray], r_grid=R_GRID, floor_rel: float = FLOOR_REL,
             max_points: int = 400_000_000) -> np.ndarray:
    """LOO log-likelihood (GPU KDE) of every dataset at h = r * thumb for each
    r in r_grid.  Returns (len(datasets), len(r_grid)) float64."""
    out = np.empty((len(datasets), len(r_grid)))
    i = 0
    while i < len(datasets):
        # group datasets so that points x candidates stay under max_points
        j, pts = i, 0
        while j < len(datasets) and (pts == 0 or pts + datasets[j].shape[1] * len(r_grid) <= max_points):
            pts += datasets[j].shape[1] * len(r_grid)
            j += 1
        group = datasets[i:j]
        flat = np.concatenate([g.reshape(-1) for g in group]).astype(np.float32)
        dptr = ctx.alloc(flat.nbytes)
        try:
            ctx.h2d(dptr, flat)
            specs, off = [], 0
            for g in group:
                d, n = g.shape
                th = thumb(g)
                th[th <= 0] = 1.0
                for r in r_grid:
                    specs.append({"n_dim": d, "n": n, "n_rows": n, "off": [off + e * n for e in range(d)],
                                  "h": [float(r * th[e]) for e in range(d)]})
                off += d * n
            res = batch_loo(ctx, dptr, specs, floor_rel)
        finally:
            ctx.free(dptr)
        out[i:j] = res.reshape(len(group), len(r_grid))
        i = j
    return out


def label_from_grid(ll: np.ndarray, r_grid=R_GRID) -> float:
    """log2 of the LOO-optimal multiplier, refined by a parabola through the
    best grid point and its neighbours (in log2 r)."""
    lr = np.log2(r_grid)
    k = int(np.argmax(ll))
    if 0 < k < len(ll) - 1:
        y0, y1, y2 = ll[k - 1], ll[k], ll[k + 1]
        den = y0 - 2 * y1 + y2
        if den < 0:
            step = lr[1] - lr[0]
            return float(lr[k] + 0.5 * step * (y0 - y2) / den)
    return float(lr[k])


# ---------------------------------------------------------------- model
class BandwidthModel:
    """The MLP on the device (ent_mlp) plus its input normalisation."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        check(load().ent_mlp_create(ctx.handle, N_FEATURES, HIDDEN, HIDDEN_LAYERS, seed, C.byref(h)))
        self._h = h
        n = C.c_size_t()
        check(load().ent_mlp_num_params(h, C.byref(n)))
        self.n_params = n.value
        self.mean = np.zeros(N_FEATURES, np.float32)
        self.std = np.ones(N_FEATURES, np.float32)
        # per-feature range of the training set (the network's envelope)
        self.fmin = np.full(N_FEATURES, -np.inf, np.float32)
        self.fmax = np.full(N_FEATURES, np.inf, np.float32)

    def close(self) -> None:
        if self._h is not None:
            load().ent_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ---------------------------------------------------------
    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        check(load().ent_mlp_get_params(self.ctx.handle, self._h, _ptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(load().ent_mlp_set_params(self.ctx.handle, self._h, _ptr(p)))

    def save(self, path: Path, meta: Optional[dict] = None) -> None:
        np.savez_compressed(path, params=self.get_params(), mean=self.mean, std=self.std, fmin=self.fmin,
                            fmax=self.fmax, meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, ctx: Context, path: Path = WEIGHTS) -> "BandwidthModel":
        z = np.load(path, allow_pickle=False)
        m = cls(ctx)
        m.set_params(z["params"])
        m.mean = z["mean"].astype(np.float32)
        m.std = z["std"].astype(np.float32)
        if "fmin" in z.files:
            m.fmin = z["fmin"].astype(np.float32)
            m.fmax = z["fmax"].astype(np.float32)
        return m

    # compute -----------------------------------------------------------
    def _norm(self, F: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(((np.asarray(F, np.float32) - self.mean) / self.std).astype(np.float32))

    def predict_raw(self, Xn: np.ndarray) -> np.ndarray:
        Xn = np.ascontiguousarray(Xn, dtype=np.float32)
        y = np.empty(Xn.shape[0], np.float32)
        check(load().ent_mlp_predict(self.ctx.handle, self._h, _ptr(Xn), Xn.shape[0], _ptr(y)))
        return y

    def gradient_raw(self, Xn: np.ndarray, t: np.ndarray):
        Xn = np.ascontiguousarray(Xn, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        check(load().ent_mlp_gradient(self.ctx.handle, self._h, _ptr(Xn), _ptr(t), Xn.shape[0], _ptr(g),
                                      C.byref(loss)))
        return g, loss.value

    def train_raw(self, Xn: np.ndarray, t: np.ndarray, epochs: int, batch: int, lr: float, seed: int = 0):
        Xn = np.ascontiguousarray(Xn, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n = Xn.shape[0]
        rng = np.random.default_rng(seed)
        order = np.concatenate([rng.permutation(n) for _ in range(epochs)]).astype(np.uint32)
        losses = np.empty(epochs, np.float32)
        check(load().ent_mlp_train(self.ctx.handle, self._h, _ptr(Xn), _ptr(t), n, _ptr(order), epochs, batch,
                                   float(lr), _ptr(losses)))
        return losses

    def fit(self, F: np.ndarray, t: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 1e-3,
            seed: int = 0) -> np.ndarray:
        F = np.asarray(F, np.float32)
        self.fmin = F.min(axis=0)
        self.fmax = F.max(axis=0)
        self.mean = F.mean(axis=0)
        self.std = F.std(axis=0)
        self.std[self.std < 1e-6] = 1.0
        return self.train_raw(self._norm(F), t, epochs, batch, lr, seed)

    def predict_log2r(self, F: np.ndarray) -> np.ndarray:
        """log2 of the multiplier of the rule of thumb.  A sample larger (or
        smaller) than the training range is evaluated at the nearest trained
        N, and the multiplier then follows the asymptotic scaling of the
        AMISE-optimal bandwidth, N^(-1/(d+4)) against the rule of thumb's
        N^(-1/7): feature 22 is that log ratio, so the correction is its
        change between the two N."""
        F = np.array(np.atleast_2d(F), dtype=np.float64)
        lo_n, hi_n = float(self.fmin[0]), float(self.fmax[0])
        log10n = F[:, 0].copy()
        clamped = np.clip(log10n, lo_n, hi_n) if np.isfinite(lo_n) and np.isfinite(hi_n) else log10n
        d = F[:, 1]
        def f22(l10):
            n = 10.0 ** l10
            return np.log2(n ** (-1.0 / (d + 4.0)) / n ** (-1.0 / 7.0))
        corr = f22(log10n) - f22(clamped)
        F[:, 0] = clamped
        F[:, 22] = f22(clamped)
        lr = self.predict_raw(self._norm(F.astype(np.float32))) + corr
        lo, hi = np.log2(R_GRID[0]), np.log2(R_GRID[-1])
        return np.clip(lr, lo, hi)

    # features that depend on N only (log10 N, the log ratio of the N
    # exponents): outside the envelope check
    N_FEATURES_OF_N = (0, 22)

    def in_envelope(self, F: np.ndarray, margin: float = 0.05) -> np.ndarray:
        """True per row of F whose shape features all lie within the training
        range (widened by `margin` of each feature's span); the features of
        N alone are not checked."""
        F = np.atleast_2d(np.asarray(F, np.float32))
        span = self.fmax - self.fmin
        span = np.where(np.isfinite(span), span, 0.0)
        ok = (F >= self.fmin - margin * span) & (F <= self.fmax + margin * span)
        ok[:, list(self.N_FEATURES_OF_N)] = True
        return np.all(ok, axis=1)

    def bandwidth(self, x: np.ndarray) -> np.ndarray:
        """Per-dimension bandwidth for a (d, N) sample (no envelope check)."""
        return thumb(x) * 2.0 ** float(self.predict_log2r(features(x))[0])


def column_bandwidths(ctx: Context, model: BandwidthModel, x_dev, X: Optional[np.ndarray] = None):
    """Learned bandwidth of every column of the device array x_dev (n_cols,
    n_rows): h = r * thumb, thumb = N^(-1/7) * sd (hestimate.cpp:19-30), with
    the network's r where the column's features lie in the training
    envelope, else the AMISE bandwidth (hestimate, computed on the device).
    The features' statistics and the network run on the device.  (X, the
    host copy, is not needed.)  Returns (h float32 (n_cols,), used_network
    bool)."""
    from .kde import hestimate

    n_cols, n_rows = x_dev.shape
    F, st = column_features_dev(ctx, x_dev.ptr, n_rows, n_cols)
    th = (float(n_rows) ** (-1.0 / 7.0)) * np.sqrt(st[:, 1])
    ok = model.in_envelope(F)
    h = (th * 2.0 ** model.predict_log2r(F)).astype(np.float32)
    if not ok.all():
        h[~ok] = hestimate(ctx, x_dev, "amise")[~ok]
    return h, ok


# -------------------------------------------------------------- training
def make_training_set(ctx: Context, n_datasets: int, seed: int, log=print):
    rng = np.random.default_rng(seed)
    F = np.empty((n_datasets, N_FEATURES), np.float32)
    Y = np.empty(n_datasets, np.float32)
    LL = np.empty((n_datasets, len(R_GRID)))
    dims = np.empty(n_datasets, np.int32)
    chunk = 500
    t0 = time.time()
    for s in range(0, n_datasets, chunk):
        ds = [synth_dataset(rng) for _ in range(min(chunk, n_datasets - s))]
        ll = loo_grid(ctx, ds)
        for k, (x, row) in enumerate(zip(ds, ll)):
            F[s + k] = features(x)
            Y[s + k] = label_from_grid(row)
            LL[s + k] = row
            dims[s + k] = x.shape[0]
        log(f"[learned_h] labelled {s + len(ds)}/{n_datasets} datasets ({time.time() - t0:.1f} s)")
    return F, Y, LL, dims


def evaluate_mi(ctx: Context, model: BandwidthModel, n_cols: int = 64, n_rows: int = 200_000, log=print) -> dict:
    """The learned bandwidth on the MI workload's columns (64 x 200,000
    mixed columns, datasets.mixed_columns) against hestimate: per column the
    rule of thumb, the AMISE bandwidth and the network's choice (with its
    AMISE fallback); each scored by the entropy it gives relative to the
    AMISE bandwidth's entropy, by its leave-one-out log-likelihood per point
    against the best of the 32-multiplier grid, and by the MI matrix it
    gives.  All densities on the GPU."""
    from . import datasets
    from .kde import DeviceArray, batch_entropy, hestimate, mi_matrix

    X = datasets.mixed_columns(n_cols, n_rows)
    dx = DeviceArray(ctx, X)
    try:
        t0 = time.perf_counter()
        h_thumb = hestimate(ctx, dx, "thumb")
        t_thumb = time.perf_counter() - t0
        t0 = time.perf_counter()
        h_amise = hestimate(ctx, dx, "amise")
        t_amise = time.perf_counter() - t0
        t0 = time.perf_counter()
        h_mlp, used = column_bandwidths(ctx, model, dx, X)
        t_mlp = time.perf_counter() - t0
        grid = loo_grid(ctx, [X[c: c + 1] for c in range(n_cols)]) / n_rows  # (n_cols, 32) per point
        best = grid.max(axis=1)
        out = {"n_cols": n_cols, "n_rows": n_rows, "network_used_columns": int(used.sum()),
               "seconds": {"thumb": t_thumb, "amise": t_amise, "network": t_mlp}}
        H, LL, MI = {}, {}, {}
        for name, h in (("thumb", h_thumb), ("amise", h_amise), ("network", h_mlp)):
            specs = [{"n_dim": 1, "n": n_rows, "n_rows": n_rows, "off": [c * n_rows], "h": [float(h[c])]}
                     for c in range(n_cols)]
            H[name] = batch_entropy(ctx, dx.ptr, specs)
            LL[name] = batch_loo(ctx, dx.ptr, specs) / n_rows
            MI[name] = mi_matrix(ctx, dx, [float(v) for v in h])
        iu = np.triu_indices(n_cols, 1)
        for name in ("thumb", "network"):
            rel = np.abs(H[name] - H["amise"]) / np.abs(H["amise"])
            dmi = np.abs(MI[name][iu] - MI["amise"][iu])
            out[name] = {
                "entropy_rel_err_vs_amise_mean": float(rel.mean()),
                "entropy_rel_err_vs_amise_max": float(rel.max()),
                "mi_abs_diff_vs_amise_mean": float(dmi.mean()),
                "mi_rel_diff_vs_amise_mean": float(dmi.mean() / np.abs(MI["amise"][iu]).mean()),
                "loo_per_point_mean": float(LL[name].mean()),
                "loo_gap_to_grid_best_mean": float((best - LL[name]).mean()),
                "frac_columns_loo_above_amise": float(np.mean(LL[name] > LL["amise"])),
                "h_over_amise_median": float(np.median(h_thumb / h_amise if name == "thumb" else h_mlp / h_amise)),
            }
        out["amise"] = {"loo_per_point_mean": float(LL["amise"].mean()),
                        "loo_gap_to_grid_best_mean": float((best - LL["amise"]).mean())}
        out["grid_best_loo_per_point_mean"] = float(best.mean())
        log(f"[learned_h] MI workload: {json.dumps(out)}")
        return out
    finally:
        dx.free()


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="train the learned bandwidth model on the GPU")
    ap.add_argument("--datasets", type=int, default=20000)
    ap.add_argument("--holdout", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=0xE17A0005)
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    ap.add_argument("--report", type=Path, default=None)
    ap.add_argument("--eval-mi-only", action="store_true",
                    help="skip training; evaluate the weights at --out on the MI workload")
    ap.add_argument("--no-eval-mi", action="store_true")
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.eval_mi_only:
        model = BandwidthModel.load(ctx, a.out)
        rep = json.loads(a.report.read_text()) if a.report and a.report.exists() else {}
        rep["mi_workload"] = evaluate_mi(ctx, model)
        print(json.dumps(rep["mi_workload"]), flush=True)
        if a.report:
            a.report.write_text(json.dumps(rep, indent=1))
        return 0
    t0 = time.time()
    F, Y, LL, dims = make_training_set(ctx, a.datasets + a.holdout, a.seed)
    label_s = time.time() - t0
    ntr = a.datasets
    model = BandwidthModel(ctx, seed=1)
    t0 = time.time()
    losses = model.fit(F[:ntr], Y[:ntr], epochs=a.epochs)
    train_s = time.time() - t0
    pred = model.predict_log2r(F[ntr:])
    lr = np.log2(R_GRID)
    ll_at = lambda row, l2r: np.interp(l2r, lr, row)
    gain_mlp = np.array([ll_at(LL[ntr + k], pred[k]) for k in range(a.holdout)])
    gain_thumb = np.array([ll_at(LL[ntr + k], 0.0) for k in range(a.holdout)])
    best = LL[ntr:].max(axis=1)
    n_pts = 10.0 ** F[ntr:, 0].astype(np.float64)
    rep = {
        "datasets": a.datasets, "holdout": a.holdout, "epochs": a.epochs, "n_range": [N_MIN, N_MAX],
        "label_s": label_s, "train_s": train_s,
        "final_train_mse": float(losses[-1]), "first_train_mse": float(losses[0]),
        "holdout_mae_log2r_mlp": float(np.mean(np.abs(pred - Y[ntr:]))),
        "holdout_mae_log2r_thumb": float(np.mean(np.abs(Y[ntr:]))),
        "holdout_mean_loo_gap_mlp": float(np.mean(best - gain_mlp)),
        "holdout_mean_loo_gap_thumb": float(np.mean(best - gain_thumb)),
        "holdout_mean_loo_gap_per_point_mlp": float(np.mean((best - gain_mlp) / n_pts)),
        "holdout_mean_loo_gap_per_point_thumb": float(np.mean((best - gain_thumb) / n_pts)),
        "holdout_frac_mlp_better": float(np.mean(gain_mlp > gain_thumb)),
    }
    model.save(a.out, meta=rep)
    print(json.dumps(rep), flush=True)
    if not a.no_eval_mi:
        rep["mi_workload"] = evaluate_mi(ctx, model)
    if a.report:
        a.report.write_text(json.dumps(rep, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
