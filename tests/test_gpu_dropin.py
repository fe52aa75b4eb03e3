"""The C++ drop-in (include/entropical/densities.hpp) called the way the
reference drivers call it (tests/cpp/dropin_check.cpp), against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "dropin_check"


def test_cpp_dropin(tmp_path, orc):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    r = np.random.default_rng(42)
    n, c, h = 2000, 3, 0.3
    x = r.standard_normal((c, n)).astype(np.float32)
    x[1] += 0.5 * x[0]
    f = tmp_path / "x.f32"
    x.tofile(f)
    p = subprocess.run([str(EXE), str(f), str(n), str(c), str(h), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rd = lambda name: np.fromfile(tmp_path / name, dtype=np.float64)
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    assert rel(rd("p0.f64"), orc.density_fp64(x, n, [0], [h])) < 1e-5
    assert rel(rd("p01.f64"), orc.density_fp64(x, n, [0, 1], [h, h])) < 1e-5
    assert rel(rd("p012_lag.f64"), orc.density_fp64(x, n, [0, 1, 2], [h] * 3, [0, 2, 1])) < 1e-5
    # Mi::main's loop (FP32 accumulation, as the reference) vs the same
    # formula over FP64 oracle densities
    mi = rd("mi.f64").reshape(c, c)
    p1 = [orc.density_fp64(x, n, [i], [h]) for i in range(c)]
    for i in range(c):
        for j in range(i + 1, c):
            ref = orc.mi_ref(orc.density_fp64(x, n, [i, j], [h, h]), p1[i], p1[j])
            assert mi[i, j] == pytest.approx(ref, rel=1e-4, abs=1e-3)
    xs = np.sort(x[0])
    assert rd("epa.f64")[0] == pytest.approx(orc.epa_conv_fp64(xs, h), rel=1e-5)
    np.testing.assert_array_equal(rd("prep.f64").astype(np.float32), orc.coord_prep(x, n, [1, 0], [3, 0], True))
