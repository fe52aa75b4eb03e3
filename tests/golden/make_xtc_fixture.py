"""Writes the XTC fixtures of tests/test_xtc.py with the reference's own
XTC writer (the vendored xdrfile library, oracle/_ref/libref_xdr.so, built by
`make -C oracle ref`) and stores, next to each file, the coordinates the
reference's reader returns for it.

    python tests/golden/make_xtc_fixture.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
import oracle  # noqa: E402


def water_like(rng, n_frames, n_mol):
    """Triplets of nearby atoms (runs of small differences) spread in a box."""
    centres = rng.uniform(0.0, 5.0, size=(n_frames, n_mol, 1, 3))
    offs = rng.normal(0.0, 0.08, size=(n_frames, n_mol, 3, 3))
    return (centres + offs).reshape(n_frames, 3 * n_mol, 3).astype(np.float32)


def main():
    rng = np.random.default_rng(0xE17A0006)
    walk = np.cumsum(rng.normal(0.0, 0.05, size=(400, 10, 3)), axis=0) + rng.uniform(0, 3, size=(1, 10, 3))
    cases = {
        "xtc_water_3x60": (water_like(rng, 3, 20), 1000.0),
        "xtc_few_atoms_4x5": (rng.uniform(-3, 3, size=(4, 5, 3)).astype(np.float32), 1000.0),
        # a short trajectory for the command line's .xtc input (400 frames, 10 atoms)
        "xtc_walk_400x10": (walk.astype(np.float32), 1000.0),
    }
    for name, (xyz, prec) in cases.items():
        path = HERE / f"{name}.xtc"
        oracle.ref_write_xtc(path, xyz, prec)
        back = oracle.ref_read_xtc(path, xyz.shape[0] + 1, xyz.shape[1])
        np.save(HERE / f"{name}.expected.npy", back)
        print(path, back.shape)


if __name__ == "__main__":
    main()
