"""XTC trajectory reader (ent_read_xtc, host code of the library) against the
reference's own XTC writer and reader (the vendored xdrfile library compiled
into oracle/_ref): committed fixtures written by the reference, and
randomised round trips through every branch of the format (plain floats for
<= 9 atoms, mixed-radix first atoms, three-field first atoms for ranges over
2^24, runs of small differences with smallidx moving both ways).
Comparison is bit for bit."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def _as_frames(cols, natoms):
    # (3*natoms, n_frames) column-major -> (n_frames, natoms, 3)
    return cols.T.reshape(-1, natoms, 3)


@pytest.mark.parametrize("name,natoms", [("xtc_water_3x60", 60), ("xtc_few_atoms_4x5", 5), ("xtc_walk_400x10", 10)])
def test_golden_fixtures(name, natoms):
    from entropical_amd import read_xtc

    got = read_xtc(GOLD / f"{name}.xtc")
    exp = np.load(GOLD / f"{name}.expected.npy")
    assert got.shape == (3 * natoms, exp.shape[0])
    np.testing.assert_array_equal(_as_frames(got, natoms).view(np.uint32), exp.view(np.uint32))


def _cases(rng):
    def water(nf, nmol, spread, box):
        c = rng.uniform(0.0, box, size=(nf, nmol, 1, 3))
        return (c + rng.normal(0, spread, size=(nf, nmol, 3, 3))).reshape(nf, 3 * nmol, 3)

    yield "single atom", rng.normal(size=(3, 1, 3)), 1000.0
    yield "nine atoms", rng.normal(size=(2, 9, 3)), 1000.0
    yield "ten atoms", rng.normal(size=(2, 10, 3)), 1000.0
    yield "water 300", water(4, 100, 0.08, 6.0), 1000.0
    yield "water, low precision", water(3, 200, 0.05, 20.0), 100.0
    yield "scattered", rng.uniform(-50, 50, size=(3, 500, 3)), 1000.0
    yield "mixed spacing", np.concatenate([water(2, 50, 0.01, 2.0), water(2, 50, 1.0, 200.0)], axis=1), 1000.0
    yield "range over 2^24", rng.uniform(-20000, 20000, size=(2, 300, 3)), 1000.0
    yield "tight clusters", water(2, 400, 0.002, 3.0), 1000.0


def test_round_trips_against_reference_reader(tmp_path):
    import oracle
    from entropical_amd import read_xtc

    if not oracle.xdr_available():
        pytest.skip("reference XTC library not built (make -C oracle ref)")
    rng = np.random.default_rng(7)
    for k, (label, xyz, prec) in enumerate(_cases(rng)):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        path = tmp_path / f"c{k}.xtc"
        oracle.ref_write_xtc(path, xyz, prec)
        ref = oracle.ref_read_xtc(path, xyz.shape[0] + 1, xyz.shape[1])
        got = _as_frames(read_xtc(path), xyz.shape[1])
        assert got.shape == ref.shape, label
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=label)
        # and the decoded values are the written ones to the precision
        assert np.max(np.abs(got - xyz)) <= 0.51 / prec + 1e-6 * np.max(np.abs(xyz)), label


def test_errors(tmp_path):
    from entropical_amd import EntropicalError, read_xtc

    with pytest.raises(EntropicalError):
        read_xtc(tmp_path / "missing.xtc")
    data = (GOLD / "xtc_water_3x60.xtc").read_bytes()
    cut = tmp_path / "cut.xtc"
    cut.write_bytes(data[: len(data) - 40])
    with pytest.raises(EntropicalError):
        read_xtc(cut)
    bad = tmp_path / "bad.xtc"
    bad.write_bytes(b"\x00\x00\x00\x01" + data[4:])
    with pytest.raises(EntropicalError):
        read_xtc(bad)
