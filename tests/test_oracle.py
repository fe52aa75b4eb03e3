"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something other than its own
formula: closed forms, invariants that a dropped term / wrong sign / wrong
index would break, brute force vs windowed evaluation, and the reference's
own code where it compiles (oracle/_ref: prob_dens_coord_prep and
Transs::BoxedSearch::Boxes).
"""
import math

import numpy as np
import pytest


def rng(seed=0):
    return np.random.default_rng(seed)


# ------------------------------------------------------------ densities FP64
def test_single_point_closed_form(orc):
    for h in (0.1, 0.37, 2.0):
        p = orc.density_fp64(np.array([[1.25]], np.float32), 1, [0], [h])
        assert p[0] == pytest.approx(0.75 / float(np.float32(h)), rel=1e-15)


def test_two_points_closed_form(orc):
    h = 0.5
    delta = 0.25
    x = np.array([[0.0, delta]], np.float32)
    p = orc.density_fp64(x, 2, [0], [h])
    u2 = (delta / h) ** 2
    expect = (0.75 / h) * (1.0 + (1.0 - u2)) / 2.0
    np.testing.assert_allclose(p, [expect, expect], rtol=1e-15)
    # outside the support: only the self pair counts
    x = np.array([[0.0, 0.75]], np.float32)
    np.testing.assert_allclose(orc.density_fp64(x, 2, [0], [h]), [0.75 / h / 2] * 2, rtol=1e-15)


def test_support_edge_is_zero(orc):
    # |u| == 1 exactly contributes 0 (K(1) = 0); both sides of the edge agree
    h = 0.5
    x = np.array([[0.0, 0.5]], np.float32)
    np.testing.assert_allclose(orc.density_fp64(x, 2, [0], [h]), [0.75 / h / 2] * 2, rtol=1e-15)


def test_constant_column_separates(orc):
    r = rng(1)
    n = 300
    x = r.standard_normal(n).astype(np.float32)
    c = np.full(n, 3.0, np.float32)
    coords = np.stack([x, c])
    p1 = orc.density_fp64(coords, n, [0], [0.3])
    p2 = orc.density_fp64(coords, n, [0, 1], [0.3, 0.7])
    p3 = orc.density_fp64(coords, n, [0, 1, 1], [0.3, 0.7, 0.2])
    f = lambda v: float(np.float32(v))
    np.testing.assert_allclose(p2, p1 * 0.75 / f(0.7), rtol=1e-13)
    np.testing.assert_allclose(p3, p1 * (0.75 / f(0.7)) * (0.75 / f(0.2)), rtol=1e-13)


def test_dimension_order_and_bandwidth_binding(orc):
    # swapping the columns together with their bandwidths leaves p unchanged;
    # swapping only the bandwidths does not (catches a transposed h index)
    r = rng(2)
    n = 400
    coords = np.stack([r.standard_normal(n), 3 * r.standard_normal(n)]).astype(np.float32)
    a = orc.density_fp64(coords, n, [0, 1], [0.2, 0.9])
    b = orc.density_fp64(coords, n, [1, 0], [0.9, 0.2])
    c = orc.density_fp64(coords, n, [0, 1], [0.9, 0.2])
    np.testing.assert_allclose(a, b, rtol=1e-13)
    assert np.max(np.abs(a - c) / a) > 1e-2


def test_scaling_invariance(orc):
    r = rng(3)
    n = 500
    x = r.standard_normal((2, n)).astype(np.float32)
    p = orc.density_fp64(x, n, [0, 1], [0.3, 0.4])
    q = orc.density_fp64(4 * x, n, [0, 1], [1.2, 1.6])
    np.testing.assert_allclose(q, p / 16.0, rtol=1e-13)


def test_window_equals_brute(orc):
    r = rng(4)
    for d, n in ((1, 700), (2, 600), (3, 500)):
        x = r.standard_normal((3, n)).astype(np.float32)
        cols = list(range(d))
        h = [0.25, 0.4, 0.6][:d]
        a = orc.density_fp64(x, n, cols, h)
        b = orc.density_fp64_window(x, n, cols, h)
        np.testing.assert_allclose(a, b, rtol=1e-13)


def test_lagged_density_normalises_by_original_rows(orc):
    r = rng(5)
    n, tau = 300, 7
    x = r.standard_normal((1, n)).astype(np.float32)
    p = orc.density_fp64(x, n, [0, 0], [0.3, 0.3], [0, tau])
    assert p.shape == (n - tau,)
    # same points materialised explicitly, normalised by n - tau
    pts = np.stack([x[0, : n - tau], x[0, tau:]])
    q = orc.density_fp64(pts, n - tau, [0, 1], [0.3, 0.3])
    np.testing.assert_allclose(p * n, q * (n - tau), rtol=1e-13)


def test_self_term_lower_bound(orc):
    r = rng(6)
    n = 800
    x = (10 * r.standard_normal((2, n))).astype(np.float32)
    p = orc.density_fp64(x, n, [0, 1], [0.1, 0.2])
    f = lambda v: float(np.float32(v))
    assert np.all(p >= (0.75 / f(0.1)) * (0.75 / f(0.2)) / n * (1 - 1e-12))


def test_brute_force_matches_numpy_on_tiny_input(orc):
    # independent evaluation of the textbook product kernel with numpy
    r = rng(7)
    n = 40
    x = r.standard_normal((2, n)).astype(np.float32)
    h = np.array([0.7, 0.9], np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    u = (xd[:, :, None] - xd[:, None, :]) / h[:, None, None]
    k = np.where(np.abs(u) <= 1, 0.75 * (1 - u * u) / h[:, None, None], 0.0)
    expect = np.prod(k, axis=0).sum(axis=1) / n
    np.testing.assert_allclose(orc.density_fp64(x, n, [0, 1], h), expect, rtol=1e-12)


# --------------------------------------------------------- FP32 replicas
def test_cuda_replica_close_to_fp64(orc):
    r = rng(8)
    n = 3000
    x = r.standard_normal((2, n)).astype(np.float32)
    for cols, h in (([0], [0.2]), ([0, 1], [0.3, 0.25])):
        a = orc.density_cuda_replica(x, n, cols, h)
        b = orc.density_fp64(x, n, cols, h)
        rel = np.abs(a - b) / b
        assert rel.max() < 2e-5  # FP32 accumulation of ~n*support terms


def test_omp_baseline_close_to_fp64(orc):
    r = rng(9)
    n = 2000
    x = r.standard_normal((3, n)).astype(np.float32)
    for d in (1, 2, 3):
        h = [0.3, 0.35, 0.4][:d]
        a = orc.density_omp_packed(x[:d], n, h)
        b = orc.density_fp64(x, n, list(range(d)), h)
        rel = np.abs(a - b) / b
        assert rel.max() < 2e-5


# --------------------------------------------------------- coord prep / ref
def test_coord_prep_indices(orc):
    n = 6
    coords = np.arange(3 * n, dtype=np.float32)  # value = col*n + row
    out = orc.coord_prep(coords, n, [2, 0], [1, 3], row_major=True)
    ns = n - 3
    expect = np.array([[2 * n + i + 1, 0 * n + i + 3] for i in range(ns)], np.float32).reshape(-1)
    np.testing.assert_array_equal(out, expect)
    out_c = orc.coord_prep(coords, n, [2, 0], [1, 3], row_major=False)
    np.testing.assert_array_equal(out_c, expect.reshape(ns, 2).T.reshape(-1))


@pytest.mark.skipif("not __import__('oracle').ref_available()")
def test_coord_prep_matches_reference_code(orc):
    r = rng(10)
    n = 97
    coords = r.standard_normal(4 * n).astype(np.float32)
    for cols, tau in (([1], [0]), ([3, 0], [0, 5]), ([2, 2, 1], [4, 0, 9])):
        h = [0.1] * len(cols)
        for rm in (True, False):
            a = orc.coord_prep(coords, n, cols, tau, rm)
            b = orc.ref_coord_prep(coords, n, cols, h, tau, rm)
            np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------- binning
def test_bin_edges(orc):
    x = np.array([0.0, 0.1, 0.2, 0.3, 0.95, 1.0], np.float32)
    box, nb, mn, mx = orc.bin_column(x, 0.25)
    assert nb == 4 and mn == 0.0 and mx == 1.0
    # x == max on an exact multiple is clamped into the last box
    assert box[-1] == 3
    assert list(box) == [0, 0, 0, 1, 3, 3]
    # all-equal data: one box
    box, nb, _, _ = orc.bin_column(np.full(5, -2.5, np.float32), 0.1)
    assert nb == 1 and np.all(box == 0)
    # negative-only data (the reference max init breaks this case)
    box, nb, mn, mx = orc.bin_column(np.array([-3.0, -2.0, -1.5], np.float32), 0.5)
    assert mx == -1.5 and nb == 3 and list(box) == [0, 2, 2]


@pytest.mark.skipif("not __import__('oracle').ref_available()")
def test_bin_matches_reference_boxes(orc):
    r = rng(11)
    for trial in range(20):
        n = int(r.integers(2, 3000))
        x = (r.standard_normal(n) * r.uniform(0.1, 5) + 1.0).astype(np.float32)
        x[0] = abs(x[0]) + 0.5  # at least one positive value (reference max init)
        bs = np.float32(r.uniform(0.01, 1.0))
        box, nb, mn, mx = orc.bin_column(x, bs)
        q = np.float32(np.float32(mx - mn) / bs)
        if q == np.floor(q):
            continue  # x == max on a box edge is undefined in the reference
        rb, rnb = orc.ref_boxes(x.reshape(1, -1), n, 0, bs)
        assert rnb == nb
        np.testing.assert_array_equal(box.astype(np.int64), rb)


def test_counting_sort_is_stable(orc):
    keys = np.array([3, 1, 3, 0, 1, 3, 0], np.uint32)
    perm, start = orc.counting_sort(keys, 4)
    assert list(perm) == [3, 6, 1, 4, 0, 2, 5]
    assert list(start) == [0, 2, 4, 4, 7]


# --------------------------------------------------------- epa convolution
def test_epa_conv_closed_form(orc):
    h = 0.5
    assert orc.epa_conv_fp64(np.array([1.0], np.float32), h) == pytest.approx(4.5, rel=1e-15)
    delta = 0.3
    x = np.array([0.0, delta], np.float32)
    d = float(np.float32(delta)) / h
    expect = (2 * 4.5 + 2 * (4.5 - 2.25 * d)) / 4
    assert orc.epa_conv_fp64(x, h) == pytest.approx(expect, rel=1e-12)
    assert orc.epa_conv_omp(x, h) == pytest.approx(expect, rel=1e-6)


def test_epa_conv_replica_close_to_fp64(orc):
    x = np.sort(rng(12).standard_normal(1500).astype(np.float32))
    a = orc.epa_conv_omp(x, 0.2)
    b = orc.epa_conv_fp64(x, 0.2)
    assert a == pytest.approx(b, rel=1e-5)


# ------------------------------------------------------------- hestimate
def test_thumb_closed_form(orc):
    x = np.array([-1.0, 1.0], np.float32)
    assert orc.thumb(x) == pytest.approx(2.0 ** (-1 / 7), rel=1e-6)
    x = np.array([2.0, 4.0, 6.0, 8.0], np.float32)
    assert orc.thumb(x) == pytest.approx(4.0 ** (-1 / 7) * math.sqrt(5.0), rel=1e-6)


def test_amise_value_closed_form(orc):
    conv, h, n = 2.5, 0.3, 1000
    a = 0.04 / h**5 * conv
    expect = a**0.2 * 1.25 * 0.6**0.8 / n**0.8
    assert orc.amise_value(conv, h, n) == pytest.approx(expect, rel=1e-6)


def test_amise_search_brackets_minimum(orc):
    x = rng(13).standard_normal(800).astype(np.float32)
    h0 = orc.thumb(x)
    h = orc.amise(x, exact=True)
    assert 0.1 * h0 <= h <= 2 * h0
    xs = np.sort(x)

    def A(hh):
        return orc.amise_value(orc.epa_conv_fp64(xs, hh), hh, x.size)

    # the search result is not worse than a coarse grid's best by much
    grid = np.linspace(0.1 * h0, 2 * h0, 60)
    best = min(A(g) for g in grid)
    assert A(h) <= best * 1.01


# ------------------------------------------------------- driver formulas
def test_mi_independent_is_zero(orc):
    r = rng(14)
    pi = r.uniform(0.1, 1, 100)
    pj = r.uniform(0.1, 1, 100)
    assert orc.mi_fp64(pi * pj, pi, pj) == pytest.approx(0.0, abs=1e-14)
    assert orc.mi_ref(pi * pj, pi, pj) == pytest.approx(0.0, abs=1e-6)


def test_mi_closed_form(orc):
    pij = np.array([0.5, 0.25])
    pi = np.array([0.5, 0.5])
    pj = np.array([0.5, 0.5])
    expect = 0.5 * math.log(2.0) + 0.25 * math.log(1.0)
    assert orc.mi_fp64(pij, pi, pj) == pytest.approx(expect, rel=1e-15)
    assert orc.mi_ref(pij, pi, pj) == pytest.approx(expect, rel=1e-6)
    assert orc.mi_fp64(pij, pi, pj, bits=True) == pytest.approx(0.5, rel=1e-15)


def test_te_guards(orc):
    p3 = np.array([0.5, 0.0, 2.0, 0.25])
    p1 = np.array([1.0, 1.0, 1.0, 1.0])
    p2 = np.array([0.5, 0.5, 0.5, 0.0])  # inv(0) -> 0 -> log(0) -> 0 in select_log
    pt = np.array([0.5, 0.5, 0.5, 0.5])
    v, strange = orc.te_ref(p3, p1, p2, pt, pref=2.0)
    # k=0: 0.5*2*log(0.5*1/0.5/0.5)=log 2 ; k=1 skipped (p3=0); k=2 rejected (p3>1);
    # k=3: ratio 0 -> select_log(0) = 0
    assert v == pytest.approx(math.log(2.0), rel=1e-6)
    assert strange == 1


def test_te_fp64_conditional_independence_is_zero(orc):
    # p(x_t+tau, x_t, y_t) = p(x_t+tau, x_t) p(x_t, y_t) / p(x_t): the ratio
    # p3 p_i / (p_ij p_itau) is 1 at every point, so TE vanishes exactly
    r = rng(31)
    p1 = r.uniform(0.2, 0.9, 200)
    p2 = r.uniform(0.1, 0.5, 200)
    pt = r.uniform(0.1, 0.5, 200)
    p3 = p2 * pt / p1
    assert np.all(p3 <= 1.0)
    assert orc.te_fp64(p3, p1, p2, pt, pref=0.3) == pytest.approx(0.0, abs=1e-15)


def test_te_fp64_hand_computed(orc):
    # transs.cpp:75-100: sum over p3 > 0 of p3 * pref * log(p3 p_i inv(p_ij) inv(p_itau)),
    # rejected when p3 > 1 (or not finite)
    p3 = np.array([0.5, 0.3, 2.0, 0.0, 0.2])
    p1 = np.array([0.6, 0.4, 1.0, 0.7, 0.1])     # p_i: numerator
    p2 = np.array([0.25, 0.0, 1.0, 0.3, 0.4])    # p_ij: inv(0) = 0 -> ratio 0 -> term 0
    pt = np.array([0.8, 0.5, 0.1, 0.2, 0.05])    # p_itau: denominator
    h_i, h_j = 0.5, 2.0
    pref = h_i * h_i * h_j
    # row 0: ratio 0.5*0.6/(0.25*0.8) = 1.5; row 1: inv(0); row 2: p3 > 1 (would be
    # 2*log(20)); row 3: p3 = 0 skipped; row 4: ratio 0.2*0.1/(0.4*0.05) = 1.0
    expect = 0.5 * pref * math.log(1.5)
    got = orc.te_fp64(p3, p1, p2, pt, pref=pref)
    assert got == pytest.approx(expect, rel=1e-14)
    assert got > 0
    # the same sum in bits
    assert orc.te_fp64(p3, p1, p2, pt, pref=pref, bits=True) == pytest.approx(
        0.5 * pref * math.log2(1.5), rel=1e-14)
    # a flipped ratio, or p_i and p_itau exchanged, changes the value
    assert orc.te_fp64(p3, pt, p2, p1, pref=pref) != pytest.approx(expect, rel=1e-3)
    assert orc.te_fp64(p3, p2, p1, pt, pref=pref) != pytest.approx(expect, rel=1e-3)


def test_te_fp64_agrees_with_reference_arithmetic(orc):
    # te_ref (pinned by hand above) rounds the ratio to float and uses logf;
    # te_fp64 is the same sum in FP64.  Random inputs including the guard rows.
    r = rng(32)
    n = 5000
    p3 = r.uniform(0.0, 0.9, n)
    p1 = r.uniform(0.05, 1.0, n)
    p2 = r.uniform(0.05, 1.0, n)
    pt = r.uniform(0.05, 1.0, n)
    p3[::97] = 0.0
    p2[::89] = 0.0
    pt[::83] = 0.0
    p3[::101] = 1.5
    a = orc.te_fp64(p3, p1, p2, pt, pref=0.7)
    b, strange = orc.te_ref(p3, p1, p2, pt, pref=0.7)
    assert strange == np.count_nonzero(p3 > 1.0)
    assert a == pytest.approx(b, rel=1e-6)


def test_autoent_hand_computed(orc):
    # autoent.cpp:52-57: sum over p2 > 0 of p2 h h log(p1 inv(p2 h))
    h = 0.5
    p1 = np.array([0.3, 0.5, 0.2])
    p2 = np.array([0.2, 0.0, 0.8])
    # row 0: 0.2*0.25*log(0.3/0.1) ; row 1 skipped (p2 = 0); row 2: 0.8*0.25*log(0.2/0.4)
    expect = 0.2 * 0.25 * math.log(3.0) + 0.8 * 0.25 * math.log(0.5)
    got = orc.autoent_ref(p1, p2, h)
    assert got == pytest.approx(expect, rel=1e-6)
    assert orc.autoent_ref(p1, p2, h, bits=True) == pytest.approx(
        0.2 * 0.25 * math.log2(3.0) + 0.8 * 0.25 * math.log2(0.5), rel=1e-6)
    # arguments exchanged: a different sum
    assert orc.autoent_ref(p2, p1, h) != pytest.approx(expect, rel=1e-3)


def test_autoent_closed_form_zero(orc):
    # p1 = p2 h at every point: each log argument is 1
    r = rng(33)
    h = 0.25
    p2 = r.uniform(0.1, 2.0, 300)
    assert orc.autoent_ref(p2 * h, p2, h) == pytest.approx(0.0, abs=1e-6)


def test_kldiv_asymmetric_hand_computed(orc):
    # kldiv.cpp:49-51: KL[i,j] = sum p_i log(p_i / p_j), KL[j,i] = sum p_j log(p_j / p_i)
    p = np.array([0.25, 0.75])
    q = np.array([0.5, 0.5])
    kl_pq = 0.25 * math.log(0.5) + 0.75 * math.log(1.5)
    kl_qp = 0.5 * math.log(2.0) + 0.5 * math.log(2.0 / 3.0)
    a, b = orc.kldiv_ref(p, q)
    assert a == pytest.approx(kl_pq, rel=1e-6)
    assert b == pytest.approx(kl_qp, rel=1e-6)
    assert abs(kl_pq - kl_qp) > 1e-2  # the two directions differ
    a2, b2 = orc.kldiv_ref(q, p)
    assert a2 == pytest.approx(kl_qp, rel=1e-6) and b2 == pytest.approx(kl_pq, rel=1e-6)
    a, b = orc.kldiv_ref(p, q, bits=True)
    assert a == pytest.approx(kl_pq / math.log(2.0), rel=1e-6)
    assert b == pytest.approx(kl_qp / math.log(2.0), rel=1e-6)


def test_entropy_uniform(orc):
    p = np.full(10, 2.0)
    assert orc.entropy(p) == pytest.approx(-math.log(2.0), rel=1e-15)


def test_negs_and_kl(orc):
    p = np.array([0.25, 0.25, 0.5])
    assert orc.negs_ref(p) == pytest.approx(-(0.5 * math.log(0.25) + 0.5 * math.log(0.5)), rel=1e-6)
    a, b = orc.kldiv_ref(p, p)
    assert a == 0.0 and b == 0.0
    q = orc.sum1_normalized(np.array([1.0, 3.0]))
    np.testing.assert_allclose(q, [0.25, 0.75], rtol=1e-15)


def test_loo_closed_form(orc):
    # two points at distance delta: p_{-i} = K(delta) for both
    delta, h = 0.2, 0.5
    pts = np.array([0.0, delta], np.float32)
    ll = orc.loo_loglik(pts, 1, [1.0], [h], 1e-300)
    u = float(np.float32(delta)) / h
    k = 0.75 * (1 - u * u) / h
    assert ll[0] == pytest.approx(2 * math.log(k), rel=1e-12)


# ---------------------------------------------------------- MLP oracle
def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp_oracle as M

    r = rng(20)
    dims = [5, 7, 6, 1]
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = r.normal(0, 0.5, n)
    X = r.normal(size=(9, 5))
    t = r.normal(size=9)
    _, g = M.loss_and_grad(p, dims, X, t, bf16=False)
    eps = 1e-6
    for k in r.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[k] = eps
        lp, _ = M.loss_and_grad(p + e, dims, X, t, bf16=False)
        lm, _ = M.loss_and_grad(p - e, dims, X, t, bf16=False)
        assert g[k] == pytest.approx((lp - lm) / (2 * eps), rel=1e-6, abs=1e-9)


def test_bf16_rounding_known_values():
    from oracle import mlp_oracle as M

    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0], np.float32)
    # 1.00390625 = 1 + 2^-8 is a tie between 1 and 1+2^-7 -> even (1.0)
    np.testing.assert_array_equal(M.bf16_round(x), [1.0, 1.0, 1.0078125, -3.140625, 65536.0])


# ------------------------------------------------------- pairs in the support
def _count_brute(x, cols, h):
    d = np.ones((x.shape[1], x.shape[1]), dtype=bool)
    for c, hh in zip(cols, h):
        v = x[c].astype(np.float64)
        d &= np.abs(v[:, None] - v[None, :]) < float(np.float32(hh))
    return int(d.sum())


def test_count_in_support_matches_brute_force(orc):
    """Ordered pairs with a nonzero kernel term: the O(n log n) sweep equals a
    brute-force count on random data with ties, for d = 1 and 2."""
    r = np.random.default_rng(61)
    for n in (1, 2, 50, 700):
        x = r.standard_normal((3, n)).astype(np.float32)
        x[1, ::5] = 0.5  # ties in dim 1
        x[0, 1::7] = x[0, 0]  # ties in dim 0
        for cols, h in (([0], [0.3]), ([1], [0.05]), ([0, 1], [0.4, 0.2]), ([2, 0], [0.1, 1.5])):
            assert orc.count_in_support(x, n, cols, h) == _count_brute(x, cols, h), (n, cols, h)


def test_count_in_support_boundary_and_jobs(orc):
    """|x_i - x_j| = h exactly is outside the support (the kernel is zero
    there); self pairs count; the job sum adds the jobs' counts."""
    x = np.array([[0.0, 0.25, 0.5], [0.0, 0.0, 1.0]], np.float32)
    # dim 0: |0 - 0.25| = 0.25 = h -> outside; 3 self pairs only
    assert orc.count_in_support(x, 3, [0], [0.25]) == 3
    # h slightly larger: (0, .25) and (.25, .5) both ways -> 3 + 4
    assert orc.count_in_support(x, 3, [0], [0.2500001]) == 7
    # 2-D: dim 1 separates point 2 (|0 - 1| = 1 >= 0.9) -> 3 + 2
    assert orc.count_in_support(x, 3, [0, 1], [0.3, 0.9]) == 5
    r = np.random.default_rng(62)
    X = r.standard_normal((4, 300)).astype(np.float32)
    cols = [[0, 1], [2, 3], [1, 3]]
    h = [[0.2, 0.3], [0.5, 0.1], [0.25, 0.25]]
    want = sum(_count_brute(X, c, hh) for c, hh in zip(cols, h))
    assert orc.count_in_support_jobs(X, 300, cols, h) == want
