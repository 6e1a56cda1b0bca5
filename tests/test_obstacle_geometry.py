"""obstaculo's obstacle columns (3_contorno.f90:494-760), computed on the
host by the library (suli_obstacle_columns, no GPU needed), against the
columns reference builds with obst_t = 1..7, 10 wrote
(tests/golden/obstaculo/*.npz, tests/golden/gen_golden.py geometry):
integers, so bit for bit."""
import glob
import os

import numpy as np
import pytest

import pysuli

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "obstaculo", "obstaculo_*.npz")))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_columns_match_reference(path):
    d = np.load(path)
    t = int(os.path.basename(path)[len("obstaculo_"):-4])
    nx, ny, nz = [int(x) for x in d["grid"]]
    dx, dy, dz = [float(x) for x in d["h"]]
    ku, kv, kw = pysuli.obstacle_columns(t, nx, ny, nz, dx, dy, dz)
    for got, name in zip((ku, kv, kw), ("ku", "kv", "kw")):
        want = d[name]
        assert got.shape == want.shape
        assert np.array_equal(got, want.astype(np.int32)), (name, np.argwhere(got != want)[:5])


def test_geometries_have_structure():
    """Guards against a vacuous pass: across the fixtures the columns take
    several heights, including the -1 of the cut-out channel (obst_t 10)."""
    vals = set()
    for path in GOLD:
        vals |= set(np.unique(np.load(path)["kw"]).astype(int))
    assert len(GOLD) == 8 and {0, 1, 5, 13} <= vals


def test_refused_geometries():
    with pytest.raises(pysuli.SuliError, match="obst_t"):
        pysuli.obstacle_columns(8, 8, 6, 5, 0.1, 0.1, 0.1)
