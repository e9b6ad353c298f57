"""Neighbour-set oracle: brute force vs cell stencil, symmetry, FCC shells, reference fixture."""
import os

import numpy as np

from conftest import GOLDEN


def _rows(offs, idx):
    return [set(idx[offs[i]:offs[i + 1]].tolist()) for i in range(len(offs) - 1)]


def _random(n, Lb, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0, Lb[d], n) for d in range(3)]


def test_brute_equals_cells_ragged_box(orc):
    Lb = np.array([11.3, 12.9, 17.7])
    lo = np.zeros(3)
    x, y, z = _random(900, Lb, 1)
    ob, ib = orc.neighbors_brute(x, y, z, Lb, 3.3)
    oc, ic = orc.neighbors_cells(x, y, z, lo, Lb, 3.3)
    assert np.array_equal(ob, oc) and np.array_equal(ib, ic)
    assert ob[-1] > 0


def test_sets_are_symmetric_and_exclude_self(orc):
    Lb = np.array([10.0, 10.0, 10.0])
    x, y, z = _random(500, Lb, 2)
    offs, idx = orc.neighbors_brute(x, y, z, Lb, 2.8)
    rows = _rows(offs, idx)
    for i, r in enumerate(rows):
        assert i not in r
        for j in r:
            assert i in rows[j]


def test_fcc_shell_counts(orc):
    """FCC at density 0.8442 (a = 1.6796): shells of 12, 6, 24, 12, 24 lie within
    r_list = 2.8 (next shell, 8 at a sqrt(3) = 2.909, is outside): 78 per atom."""
    from copy_gpu_amd import systems

    x, y, z, Lb = systems.fcc(6, 0.8442)
    Lv = np.array([Lb] * 3)
    offs, idx = orc.neighbors_cells(x, y, z, np.zeros(3), Lv, 2.8)
    counts = np.diff(offs)
    assert np.all(counts == 78), np.unique(counts)
    # dropping the 5th shell (radius 2.6557) is visible
    offs2, _ = orc.neighbors_cells(x, y, z, np.zeros(3), Lv, 2.6)
    assert np.all(np.diff(offs2) == 54)


def test_reference_lj_salt_fixture(orc):
    """The 6000-ion example of the reference (res/json_files/LJ_salt/box_102030/rbmd.data)."""
    d = np.load(os.path.join(GOLDEN, "lj_salt_6000.npz"))
    Lb = d["hi"] - d["lo"]
    x, y, z = orc.wrap(d["x"], d["y"], d["z"], d["lo"], Lb)
    ob, ib = orc.neighbors_brute(x, y, z, Lb, 2.8)
    oc, ic = orc.neighbors_cells(x, y, z, d["lo"], Lb, 2.8)
    assert np.array_equal(ob, oc) and np.array_equal(ib, ic)
    # density 1 -> about 4/3 pi 2.8^3 = 92 neighbours per ion
    assert 80 < ob[-1] / 6000 < 100


def test_epsilon_band_of_reference_criterion(orc):
    """Two atoms at exactly r_list are excluded by r_list^2 - d2 > 1e-4, and an
    atom closer than 0.01 is excluded by d2 > 1e-4 (RunWorklet.cpp:83)."""
    Lb = np.array([20.0, 20.0, 20.0])
    x = np.array([5.0, 7.8, 5.0, 5.005, 12.0])
    y = np.array([5.0, 5.0, 7.79999, 5.0, 12.0])
    z = np.array([5.0, 5.0, 5.0, 5.0, 12.0])
    offs, idx = orc.neighbors_brute(x, y, z, Lb, 2.8)
    rows = _rows(offs, idx)
    assert 1 not in rows[0]          # exactly r_list
    assert 2 not in rows[0]          # inside by 1e-5 < band
    assert 3 not in rows[0]          # closer than 0.01
    assert rows[4] == set()
