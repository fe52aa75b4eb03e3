"""The moment identities behind the density kernel's box and cell sums,
checked in FP64 against brute-force kernel sums (no GPU).

A 1-D box with origin r and points x_j stores A = s0/h^2, B = 2 S1/h^2 and
c = s0 - S2/h^2 (s0 points, S1 and S2 the sums of e_j = x_j - r and e_j^2);
for a query x whose support holds the whole box,
    sum_j (1 - ((x - x_j)/h)^2) = c + d (B - A d),  d = x - r.
A 2-D cell with origin (r0, r1) stores N[3p + q] = s0_p s1_q sum_j e_j^p g_j^q
with s_e = (1, 2/h_e^2, -1/h_e^2); for a query whose support holds the cell,
    sum_j (1 - (dx_j/h0)^2)(1 - (dy_j/h1)^2) = sum_q b_q sum_p a_p N[3p + q],
    a = (1 - d0^2/h0^2, d0, 1), b = (1 - d1^2/h1^2, d1, 1).
Outside the support the identities do not hold (the kernel clamps at 0),
which is why only boxes and cells inside every query's support use them.
"""
import numpy as np


def epa(u):
    return np.maximum(0.0, 1.0 - u * u)


def box_terms(x, r, h):
    e = x - r
    s0, s1, s2 = len(x), e.sum(), (e * e).sum()
    return s0 / h**2, 2 * s1 / h**2, s0 - s2 / h**2


def cell_terms(x, y, r0, r1, h0, h1):
    e, g = x - r0, y - r1
    s0 = (1.0, 2.0 / h0**2, -1.0 / h0**2)
    s1 = (1.0, 2.0 / h1**2, -1.0 / h1**2)
    return np.array([s0[p] * s1[q] * np.sum(e**p * g**q) for p in range(3) for q in range(3)])


def cell_sum(N, d0, d1, h0, h1):
    a = (1 - d0 * d0 / h0**2, d0, 1.0)
    b = (1 - d1 * d1 / h1**2, d1, 1.0)
    return sum(b[q] * sum(a[p] * N[3 * p + q] for p in range(3)) for q in range(3))


def test_box_identity_inside_support():
    r = np.random.default_rng(1)
    for _ in range(200):
        h = r.uniform(0.05, 2.0)
        width = h * r.uniform(0.01, 0.5)
        org = r.uniform(-50, 50)
        x = org + r.uniform(0, width, size=int(r.integers(1, 300)))
        A, B, c = box_terms(x, org, h)
        # queries whose support holds [org, org + width]
        q = r.uniform(org + width - h, org + h, size=20)
        for xq in q:
            d = xq - org
            brute = epa((xq - x) / h).sum()
            assert abs((c + d * (B - A * d)) - brute) <= 1e-9 * max(1.0, len(x))


def test_box_identity_fails_outside_support():
    """A box reaching past a query's support: the clamped sum differs from the
    polynomial, so the interior test is what keeps the kernel exact."""
    h, org = 1.0, 0.0
    x = np.linspace(0.0, 0.5, 50)
    A, B, c = box_terms(x, org, h)
    xq = -0.8  # support [-1.8, 0.2]: points above 0.2 are outside
    brute = epa((xq - x) / h).sum()
    d = xq - org
    assert abs((c + d * (B - A * d)) - brute) > 1e-3


def test_cell_identity_inside_support():
    r = np.random.default_rng(2)
    for _ in range(200):
        h0, h1 = r.uniform(0.05, 2.0, size=2)
        w0, w1 = h0 * r.uniform(0.01, 0.5), h1 * r.uniform(0.01, 0.5)
        r0, r1 = r.uniform(-20, 20, size=2)
        n = int(r.integers(1, 200))
        x = r0 + r.uniform(0, w0, size=n)
        y = r1 + r.uniform(0, w1, size=n)
        N = cell_terms(x, y, r0, r1, h0, h1)
        for _ in range(10):
            xq = r.uniform(r0 + w0 - h0, r0 + h0)
            yq = r.uniform(r1 + w1 - h1, r1 + h1)
            brute = (epa((xq - x) / h0) * epa((yq - y) / h1)).sum()
            got = cell_sum(N, xq - r0, yq - r1, h0, h1)
            assert abs(got - brute) <= 1e-9 * max(1.0, n)


def test_cell_identity_catches_a_swapped_index():
    """A transposed moment layout (N[3q + p]) gives wrong sums for a cell with
    distinct dim-0 and dim-1 spreads."""
    r = np.random.default_rng(3)
    h0, h1 = 0.3, 0.7
    x = r.uniform(0, 0.1, size=40)
    y = r.uniform(0, 0.3, size=40)
    N = cell_terms(x, y, 0.0, 0.0, h0, h1)
    Nt = N.reshape(3, 3).T.reshape(-1)
    xq, yq = 0.05, 0.2
    brute = (epa((xq - x) / h0) * epa((yq - y) / h1)).sum()
    assert abs(cell_sum(N, xq, yq, h0, h1) - brute) < 1e-9 * 40
    assert abs(cell_sum(Nt, xq, yq, h0, h1) - brute) > 1e-3
