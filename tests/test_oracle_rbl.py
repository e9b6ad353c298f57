"""The oracle's random batch list (oracle/refcpu.c orc_rbl_forces) and its periodic-
image neighbour list, pinned without the reference running:

* the oracle's own Philox4x32-10 against the Random123 known-answer vectors;
* the shell selection keeps a pair with probability 1 / pice (binomial bound);
* special cases that reduce to the full list: pice = 1 (every shell pair kept) and a
  core radius beyond the cutoff;
* unbiasedness: the mean of the random batch forces over many list builds equals the
  full-list force within the central-limit bound (the estimator of
  ExecutionMD::ComputeRBLNearForce, rbmd/framework/run/ExecutionMD.cpp:460-566, with
  the pice_num weight of ComputeNearForceRBLWorklet, RunWorklet.cpp:1015-1022);
* Newton's third law after the mean-force correction (zero total force);
* the image-aware list equals the minimum-image list when the box is wider than two
  list radii, and lists the same atom once per image in a narrower box.
"""
import numpy as np
import pytest


def _ions(n, L, seed):
    """n +-1 ions at minimum separation 0.8 in a cube of edge L."""
    from copy_gpu_amd import systems

    x, y, z, q, _ = systems.random_ions(n, n / L ** 3, 0.8, seed=seed)
    return x, y, z, q


def test_oracle_philox_known_answers(orc):
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        assert tuple(int(v) for v in orc.philox4x32(ctr, key)) == want


def test_rbl_selection_rate(orc):
    pice, trials = 7, 70000
    kept = sum(orc.rbl_select(99, 3, i % 1000, i // 1000, pice) for i in range(trials))
    p = 1.0 / pice
    sd = np.sqrt(trials * p * (1 - p))
    assert abs(kept - trials * p) < 5 * sd
    # the selection depends on the order of the pair and on the build
    a = [orc.rbl_select(5, 1, i, i + 1, 3) for i in range(300)]
    b = [orc.rbl_select(5, 1, i + 1, i, 3) for i in range(300)]
    c = [orc.rbl_select(5, 2, i, i + 1, 3) for i in range(300)]
    assert a != b and a != c


def test_rbl_special_cases_equal_full_list(orc):
    n, L = 600, 14.0
    x, y, z, q = _ions(n, L, 3)
    Lb = np.array([L] * 3)
    lo = np.zeros(3)
    Ff, evf = orc.pair_forces(x, y, z, q, lo, Lb, 2.5, 4.0, 0.5)
    # every shell pair kept (pice 1): the full list, except pairs within 1e-4 of
    # r_core^2 or of rc^2 that neither of the reference's two lists takes
    Fr, evr = orc.rbl_forces(x, y, z, q, lo, Lb, 2.5, 4.0, 0.5, 1.7, 1, 11, 1)
    rel = np.linalg.norm(Fr - Ff) / np.linalg.norm(Ff)
    assert rel < 1e-6, rel
    assert abs(evr[1] - evf[1]) < 1e-6 * abs(evf[1])
    # the whole cutoff sphere inside the core
    Fr, _ = orc.rbl_forces(x, y, z, q, lo, Lb, 2.5, 4.0, 0.5, 3.9999, 5, 11, 1)
    assert np.linalg.norm(Fr - Ff) / np.linalg.norm(Ff) < 1e-6


def test_rbl_unbiased_and_momentum_free(orc):
    n, L, pice = 400, 12.0, 4
    x, y, z, q = _ions(n, L, 8)
    Lb = np.array([L] * 3)
    lo = np.zeros(3)
    # the expectation over the shell selection: every shell pair with weight 1 (pice 1),
    # which equals the full list up to the pairs in the 1e-4 bands around r_core^2 and
    # rc^2 that the reference's lists leave out (test above)
    Ff, evf = orc.rbl_forces(x, y, z, q, lo, Lb, 2.5, 5.0, 0.5, 2.5, 1, 2024, 0)
    builds = 400
    S = np.zeros_like(Ff)
    S2 = np.zeros_like(Ff)
    E = []
    for e in range(builds):
        F, ev = orc.rbl_forces(x, y, z, q, lo, Lb, 2.5, 5.0, 0.5, 2.5, pice, 2024, e)
        assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()  # mean removed
        S += F
        S2 += F * F
        E.append(ev[1])
    mean = S / builds
    se = np.sqrt(np.maximum(S2 / builds - mean ** 2, 1e-30) / builds)
    z = (mean - Ff) / se
    assert np.abs(z).max() < 5.5, np.abs(z).max()
    assert abs(np.mean(z)) < 5 / np.sqrt(z.size)
    # a plausible mistake -- the selected shell pairs not weighted by pice -- is far
    # outside the bound: (F(pice) - core) / pice + core has the mean of that mistake
    Fc, _ = orc.rbl_forces(x, y, z, q, lo, Lb, 2.5, 5.0, 0.5, 2.5, 1 << 16, 7, 0)
    wrong = (mean - Fc) / pice + Fc
    assert np.abs((wrong - Ff) / se).max() > 10
    e_se = np.std(E) / np.sqrt(builds)
    assert abs(np.mean(E) - evf[1]) < 5 * e_se


def test_image_list_equals_minimum_image_list_in_wide_boxes(orc):
    n, L, rl = 300, 12.0, 5.3
    x, y, z, _ = _ions(n, L, 4)
    Lb = np.array([L] * 3)
    oi, ii = orc.neighbors_images(x, y, z, Lb, rl)
    ob, ib = orc.neighbors_brute(x, y, z, Lb, rl)
    assert np.array_equal(oi, ob)
    for i in range(n):
        assert np.array_equal(np.sort(ii[oi[i]:oi[i + 1]]), ib[ob[i]:ob[i + 1]])


def test_image_list_counts_every_image_in_narrow_boxes(orc):
    """A 6-wide box with list radius 5.3: atom j appears once for each image within
    the radius, so the mean row length is the density times the sphere volume."""
    n, L, rl = 200, 6.0, 5.3
    x, y, z, _ = _ions(n, L, 5)
    Lb = np.array([L] * 3)
    oi, ii = orc.neighbors_images(x, y, z, Lb, rl)
    mean = oi[-1] / n
    expect = n / L ** 3 * 4.0 / 3.0 * np.pi * rl ** 3
    assert abs(mean - expect) < 0.05 * expect
    counts = np.bincount(ii[oi[0]:oi[1]], minlength=n)
    assert counts.max() >= 2  # some atom twice in atom 0's row
    # no image of atom i itself while the box is at least r_list wide
    assert counts[0] == 0
