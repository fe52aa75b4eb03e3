"""The entropical command-line program on the GPU: every mode's printed
result against the oracle (FP64 densities and the oracle's restatements of
the reference formulas), read back from the program's text output."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "entropical_amd" / "bin" / "entropical"
RTOL = 1e-4


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    if not EXE.exists():
        pytest.fail("CLI not built (python -m entropical_amd.build)")
    r = np.random.default_rng(31)
    n = 3000
    x = np.stack([r.standard_normal(n), 0.6 * r.standard_normal(n), r.standard_t(5, n), r.uniform(-2, 2, n)])
    x[1] += 0.7 * x[0]
    x[3] += np.roll(x[2], 1) * 0.5
    x = x.astype(np.float32)
    d = tmp_path_factory.mktemp("cli")
    coords = d / "coords.txt"
    np.savetxt(coords, x.T, fmt="%.9g")  # rows = frames, columns = series
    h = np.array([0.3, 0.25, 0.35, 0.4], dtype=np.float32)
    hfile = d / "h.txt"
    hfile.write_text(" ".join(f"{v:.9g}" for v in h) + "\n")
    return x, h, coords, hfile, d


def cli(*args):
    r = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def matrix(text):
    return np.array([[float(v) for v in line.split()] for line in text.strip().splitlines()])


def test_mi(data, orc):
    x, h, coords, hfile, _ = data
    m = matrix(cli("mi", "-i", coords, "-H", hfile))
    n = x.shape[1]
    p = [orc.density_fp64(x, n, [i], [h[i]]) for i in range(4)]
    for i in range(4):
        assert m[i, i] == 0
        for j in range(i + 1, 4):
            ref = orc.mi_fp64(orc.density_fp64(x, n, [i, j], [h[i], h[j]]), p[i], p[j], False)
            assert m[i, j] == pytest.approx(ref, rel=RTOL) and m[j, i] == m[i, j]


def test_transs(data, orc):
    x, h, coords, hfile, _ = data
    tau = 2
    m = matrix(cli("transs", "-i", coords, "-H", hfile, "-t", tau, "-B"))
    n = x.shape[1]
    f = lambda v: float(np.float32(v))
    p = [orc.density_fp64(x, n, [i], [h[i]]) for i in range(4)]
    ptau = [orc.density_fp64(x, n, [i, i], [h[i], h[i]], [0, tau]) for i in range(4)]
    k = n - tau
    for i in range(4):
        for j in range(i + 1, 4):
            pij = orc.density_fp64(x, n, [i, j], [h[i], h[j]])
            itau = orc.density_fp64(x, n, [i, i, j], [h[i], h[i], h[j]], [0, tau, 0])
            jtau = orc.density_fp64(x, n, [i, j, j], [h[i], h[j], h[j]], [0, 0, tau])
            ref_ij = orc.te_fp64(itau, p[i][:k], pij[:k], ptau[i], f(h[i]) * f(h[i]) * f(h[j]), True)
            ref_ji = orc.te_fp64(jtau, p[j][:k], pij[:k], ptau[j], f(h[i]) * f(h[j]) * f(h[i]), True)
            assert m[i, j] == pytest.approx(ref_ij, rel=RTOL, abs=1e-9)
            assert m[j, i] == pytest.approx(ref_ji, rel=RTOL, abs=1e-9)


def test_autoent_kldiv_negs(data, orc):
    x, h, coords, hfile, _ = data
    n = x.shape[1]
    ae = np.array([float(v) for v in cli("autoent", "-i", coords, "-H", hfile, "-t", 3).split()])
    kl = matrix(cli("kldiv", "-i", coords, "-H", hfile))
    ng = np.array([float(v) for v in cli("negs", "-i", coords, "-H", hfile).split()])
    pn = [orc.sum1_normalized(orc.density_fp64(x, n, [i], [h[i]])) for i in range(4)]
    for i in range(4):
        p1 = orc.density_fp64(x, n, [i], [h[i]], [3])
        p2 = orc.density_fp64(x, n, [i, i], [h[i], h[i]], [0, 3])
        assert ae[i] == pytest.approx(orc.autoent_ref(p1, p2, h[i], False), rel=RTOL)
        assert ng[i] == pytest.approx(orc.negs_ref(pn[i], False), rel=RTOL)
        for j in range(4):
            if i != j:
                assert kl[i, j] == pytest.approx(orc.kldiv_ref(pn[i], pn[j], False)[0], rel=RTOL)


def test_dens_1d_2d_and_output_file(data, orc):
    x, h, coords, hfile, d = data
    n = x.shape[1]
    out = d / "dens1.txt"
    assert cli("dens", "-i", coords, "-H", hfile, "-o", out) == ""
    got = np.loadtxt(out)
    assert got.shape == (n, 4)
    for i in range(4):
        ref = orc.sum1_normalized(orc.density_fp64(x, n, [i], [h[i]]))
        np.testing.assert_allclose(got[:, i], ref, rtol=2e-5)
    # 2-D over a column selection with lags: labels use the 1-based columns
    sel_h = d / "h_sel.txt"
    sel_h.write_text(f"{h[1]:.9g} {h[3]:.9g}\n")
    lines = cli("dens", "-i", coords, "-H", sel_h, "-c", "2 4", "-D", 2, "--taus", "0 1").strip().splitlines()
    assert len(lines) == n - 1 and all(l.split()[0] == "2_4" for l in lines)
    vals = np.array([float(l.split()[1]) for l in lines])
    ref = orc.sum1_normalized(orc.density_fp64(x, n, [1, 3], [h[1], h[3]], [0, 1]))
    np.testing.assert_allclose(vals, ref, rtol=2e-5)


def test_thumb_and_amise(data, orc):
    x, h, coords, hfile, _ = data
    th = np.array([float(v) for v in cli("thumb", "-i", coords).split()])
    am = np.array([float(v) for v in cli("amise", "-i", coords).split()])
    for i in range(4):
        assert th[i] == pytest.approx(orc.thumb(x[i]), rel=2e-6)
        h0 = orc.thumb(x[i])
        assert abs(am[i] - orc.amise(x[i], exact=True)) <= 0.05 * h0


def test_xtc_input(orc, tmp_path):
    """.xtc input (the committed fixture written by the reference's XTC
    writer): columns are atom coordinates 3a+e of each frame."""
    gold = Path(__file__).resolve().parent / "golden"
    exp = np.load(gold / "xtc_walk_400x10.expected.npy")  # (frames, atoms, 3)
    cols = exp.reshape(exp.shape[0], -1).T.astype(np.float32)  # (30, frames)
    n = cols.shape[1]
    sel = [1, 2, 6, 14]  # 1-based columns
    h = [0.3, 0.3, 0.4, 0.35]
    hfile = tmp_path / "h.txt"
    hfile.write_text(" ".join(f"{v}" for v in h))
    m = matrix(cli("mi", "-i", gold / "xtc_walk_400x10.xtc", "-c", " ".join(map(str, sel)), "-H", hfile))
    x = cols[[c - 1 for c in sel]]
    p = [orc.density_fp64(x, n, [i], [h[i]]) for i in range(4)]
    for i in range(4):
        for j in range(i + 1, 4):
            ref = orc.mi_fp64(orc.density_fp64(x, n, [i, j], [h[i], h[j]]), p[i], p[j], False)
            assert m[i, j] == pytest.approx(ref, rel=RTOL)
