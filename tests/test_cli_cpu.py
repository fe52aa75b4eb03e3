"""The entropical command-line program (entropical_amd/bin/entropical):
argument handling and input errors, which the program reports before it
touches a GPU, with the reference's messages and exit status
(entropical.cpp:84-99,226-239; tools.hxx:196-205,248-253,296-301;
dens.cpp:115-119,136-141)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

EXE = Path(__file__).resolve().parents[1] / "entropical_amd" / "bin" / "entropical"

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="CLI not built (python -m entropical_amd.build)")


def run(*args):
    r = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture
def files(tmp_path):
    x = np.random.default_rng(0).standard_normal((50, 3)).astype(np.float32)
    coords = tmp_path / "coords.txt"
    np.savetxt(coords, x, fmt="%.9g")
    hfile = tmp_path / "h.txt"
    hfile.write_text("0.3 0.3 0.3\n")
    h2 = tmp_path / "h2.txt"
    h2.write_text("0.3 0.3\n")
    return coords, hfile, h2


def test_usage_and_modes():
    rc, out, err = run()
    assert rc == 1 and "modes:" in err
    rc, out, err = run("nonsense", "-h")
    assert rc == 1 and "error: unrecognized mode 'nonsense'" in err


@pytest.mark.parametrize("mode", ["transs", "autoent", "kldiv", "mi", "dens", "negs", "amise", "thumb"])
def test_help_per_mode(mode):
    rc, out, err = run(mode, "-h")
    assert rc == 0 and out.startswith(mode + " - ") and "--input" in out


def test_argument_errors(files):
    coords, hfile, _ = files
    rc, _, err = run("mi", "-H", str(hfile))
    assert rc == 1 and "error parsing arguments!" in err
    rc, _, err = run("mi", "-i", str(coords), "--nonsense")
    assert rc == 1 and "unrecognised option" in err
    rc, _, err = run("mi", "-i", str(coords), "-H")
    assert rc == 1 and "missing" in err


def test_input_errors(files):
    coords, hfile, h2 = files
    rc, _, err = run("mi", "-i", str(coords), "-H", str(h2))
    assert rc == 1 and "error: number of given bandwidth values does not match the number of selected columns!" in err
    rc, _, err = run("mi", "-i", str(coords), "-H", str(hfile), "-c", "1 0 2")
    assert rc == 1 and "error: specify column in range of [1, N_COLUMNS]!" in err
    rc, _, err = run("mi", "-i", str(coords), "-H", str(hfile), "-c", "1 2 7")
    assert rc == 1 and "error: given column index is not compatible with available number of columns (3)" in err
    rc, _, err = run("dens", "-i", str(coords), "-H", str(hfile), "-D", "4")
    assert rc == 1 and "error: only allowed kernel dimensions are 1, 2 or 3." in err
    rc, _, err = run("dens", "-i", str(coords), "-H", str(hfile), "--taus", "1")
    assert rc == 1 and "does not make sense" in err
    rc, _, err = run("dens", "-i", str(coords), "-H", str(hfile), "-D", "2", "--taus", "1 2 3")
    assert rc == 1 and "error: number of lagtime parameters does not match specified kernel dimension!" in err


def test_ragged_input_rejected(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("1 2 3\n4 5\n")
    h = tmp_path / "h.txt"
    h.write_text("0.1 0.1 0.1")
    rc, _, err = run("mi", "-i", str(f), "-H", str(h))
    assert rc == 1 and "row 2" in err
