"""C-ABI checks that need no GPU: the library loads, exports every symbol
declared in include/*.h, and its host-only entry points behave as the
reference's (no compute call reaches a device here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(ent_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from entropical_amd import _lib

    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers each of them
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_version_and_error_text():
    from entropical_amd import _lib

    lib = _lib.load()
    assert lib.ent_version() >= 1
    assert isinstance(lib.ent_last_error(), bytes)


def test_coord_prep_matches_oracle(orc):
    from entropical_amd import prob_dens_coord_prep

    r = np.random.default_rng(0)
    n = 211
    coords = r.standard_normal((4, n)).astype(np.float32)
    for cols, tau in (([1], [0]), ([3, 0], [0, 5]), ([2, 2, 1], [4, 0, 9])):
        for rm in (True, False):
            a = prob_dens_coord_prep(coords, n, cols, [0.1] * len(cols), tau, rm)
            b = orc.coord_prep(coords, n, cols, tau, rm)
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize(
    "cols,h,tau,msg",
    [
        ([], [], None, "1, 2 or 3 dimensions"),
        ([0, 1, 2, 3], [1, 1, 1, 1], None, "1, 2 or 3 dimensions"),
        ([0, 1], [0.1], None, "bandwidth"),
        ([0, 1], [0.1, 0.1], [0], "tau"),
    ],
)
def test_argument_errors(cols, h, tau, msg):
    from entropical_amd import EntropicalError, combined_densities

    coords = np.zeros((4, 10), np.float32)
    with pytest.raises(EntropicalError, match=msg):
        combined_densities(coords, 10, cols, h, tau)


def test_coord_prep_rejects_lag_beyond_rows():
    from entropical_amd import EntropicalError, prob_dens_coord_prep

    with pytest.raises(EntropicalError, match="tau"):
        prob_dens_coord_prep(np.zeros(10, np.float32), 10, [0], [0.1], [10], True)


def test_no_device_is_a_clean_error():
    from entropical_amd import _lib

    lib = _lib.load()
    n = C.c_int(-1)
    assert lib.ent_device_count(C.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = lib.ent_ctx_create(0, C.byref(h))
    assert rc == _lib.ENT_ERR_NO_DEVICE
    assert b"no HIP-capable GPU" in lib.ent_last_error()
