"""Host-side learned-proposal tables and sampler (CPU; the library's host functions).

The sampler's distribution and weights are checked against the mixture they
define (chi-square over the lattice box plus the outside mass), unbiasedness of
the weighted estimator against the exact lattice sum, and the variance proxy
against its closed forms.
"""
import math

import numpy as np
import pytest
from scipy import stats

from copy_gpu_amd import _lib as L
from copy_gpu_amd import proposal as PR

ALPHA = 0.5
BOX = (11.0, 12.0, 13.0)


def _analytic_S(alpha, box):
    # S = sum over all m != 0 of a(m), from the per-axis theta sums
    tot = 1.0
    for Ld in box:
        m = np.arange(-60, 61)
        tot *= np.exp(-(2 * np.pi * m / Ld) ** 2 / (4 * alpha)).sum()
    return tot - 1.0


def _density(Mg, rng):
    m = PR.lattice_grid(Mg)
    a = PR.analytic_weight(m, BOX, ALPHA)
    d = a * rng.uniform(0.2, 3.0, len(m))
    return m, a, d


def test_grid_order_and_half_space():
    m = PR.lattice_grid(2)
    assert m.shape == (125, 3)
    assert (m[0] == [-2, -2, -2]).all() and (m[1] == [-1, -2, -2]).all()
    assert (m[5] == [-2, -1, -2]).all()
    hs = PR.half_space(m)
    assert hs.sum() == 62
    G = len(m)
    # -m of index i is G - 1 - i; exactly one of each +/- pair is in the half space
    assert (m[::-1] == -m).all()
    nz = ~(m == 0).all(1)
    assert (hs[nz] ^ hs[::-1][nz]).all()


@pytest.mark.parametrize("lam", [0.9, 0.5])
def test_mixture_distribution_and_weights(lam):
    rng = np.random.default_rng(11)
    Mg = 3
    m, a, d = _density(Mg, rng)
    prop = PR.Proposal(Mg, d, lam, ALPHA, BOX)
    S = _analytic_S(ALPHA, BOX)
    nz = ~(m == 0).all(1)
    q = np.where(nz, lam * np.where(nz, d, 0) / d[nz].sum() + (1 - lam) * a / S, 0.0)
    P = 4000
    draws, ws = [], []
    for step in range(40):
        mm, w = prop.sample_host(P, seed=5, step=step)
        draws.append(mm)
        ws.append(w)
    mm = np.concatenate(draws)
    w = np.concatenate(ws)
    inside = (np.abs(mm) <= Mg).all(1)
    side = 2 * Mg + 1
    idx = (mm[inside, 0] + Mg) + side * ((mm[inside, 1] + Mg) + side * (mm[inside, 2] + Mg))
    # weights: w P q(m) = a(m) inside, S / (1 - lam) outside
    assert np.allclose(w[inside] * P * q[idx], a[idx], rtol=1e-12)
    assert np.allclose(w[~inside], S / ((1 - lam) * P), rtol=1e-15)
    assert not (mm == 0).all(1).any()
    # chi-square over the box cells plus the outside bin
    cnt = np.bincount(idx, minlength=len(m))[nz]
    exp = q[nz] * len(mm)
    obs = np.append(cnt, (~inside).sum())
    expv = np.append(exp, len(mm) - exp.sum())
    keep = expv > 5
    chi2 = ((obs[keep] - expv[keep]) ** 2 / expv[keep]).sum()
    dof = keep.sum() - 1
    assert stats.chi2.sf(chi2, dof) > 1e-4


def test_weighted_estimator_unbiased():
    rng = np.random.default_rng(2)
    Mg = 4
    m, a, d = _density(Mg, rng)
    prop = PR.Proposal(Mg, d, 0.9, ALPHA, BOX)
    r0 = np.array([1.3, -0.4, 2.2])

    def g(mm):
        k = PR.kvecs(mm, BOX)
        k2 = (k * k).sum(1)
        return np.cos(k @ r0) / k2

    mg = np.stack(np.meshgrid(*[np.arange(-25, 26)] * 3, indexing="ij"), -1).reshape(-1, 3)
    mg = mg[~(mg == 0).all(1)]
    exact = (PR.analytic_weight(mg, BOX, ALPHA) * g(mg)).sum()
    P = 2000
    est = np.array([(w * g(mm)).sum() for mm, w in
                    (prop.sample_host(P, seed=9, step=s) for s in range(200))])
    se = est.std(ddof=1) / math.sqrt(len(est))
    assert abs(est.mean() - exact) < 4.5 * se


def test_host_sampler_deterministic_and_seeded():
    rng = np.random.default_rng(4)
    m, a, d = _density(3, rng)
    prop = PR.Proposal(3, d, 0.9, ALPHA, BOX)
    m1, w1 = prop.sample_host(64, 1, 7)
    m2, w2 = prop.sample_host(64, 1, 7)
    m3, _ = prop.sample_host(64, 1, 8)
    assert (m1 == m2).all() and (w1 == w2).all()
    assert not (m1 == m3).all()


def test_create_rejects_bad_input():
    lib = L.lib()
    import ctypes as C
    Mg = 2
    G = (2 * Mg + 1) ** 3
    box = (C.c_double * 3)(*BOX)
    h = C.c_void_p()
    good = np.ones(G)
    for dens, lam, alpha in [(-good, 0.9, ALPHA), (np.zeros(G), 0.9, ALPHA),
                             (good * np.nan, 0.9, ALPHA), (good, 1.0, ALPHA),
                             (good, 0.0, ALPHA), (good, 0.9, 0.0)]:
        dd = np.ascontiguousarray(dens)
        rc = lib.rbmd_proposal_create(Mg, dd.ctypes.data_as(C.c_void_p), lam, alpha, box,
                                      C.byref(h))
        assert rc == L.RBMD_ERR_INVALID
    # density only at m = 0 is all zero after the origin is dropped
    z = np.zeros(G)
    z[G // 2] = 1.0
    rc = lib.rbmd_proposal_create(Mg, z.ctypes.data_as(C.c_void_p), 0.9, ALPHA, box, C.byref(h))
    assert rc == L.RBMD_ERR_INVALID


def test_variance_factor_closed_forms():
    Mg = 6
    m = PR.lattice_grid(Mg)
    k = PR.kvecs(m, BOX)
    k2 = (k * k).sum(1)
    a = PR.analytic_weight(m, BOX, ALPHA)
    nz = k2 > 0
    # |rho|^2 = c k^2 makes t constant: the analytic proposal is already optimal
    S2 = 3.0 * k2
    base = PR.variance_factor(S2, None, BOX, ALPHA, Mg)
    assert PR.variance_factor(S2, a, BOX, ALPHA, Mg, lam=1.0) == pytest.approx(base, rel=1e-12)
    assert base == pytest.approx(3.0 * a[nz].sum() ** 2, rel=1e-12)
    # general spectrum: q ~ a t attains (sum a t)^2, the Cauchy-Schwarz minimum
    S2 = np.where(nz, 1.0 + 0.5 * np.sin(k2), 0.0)
    t = np.sqrt(S2[nz] / k2[nz])
    opt = np.zeros(len(m))
    opt[nz] = a[nz] * t
    vopt = PR.variance_factor(S2, opt, BOX, ALPHA, Mg, lam=1.0)
    assert vopt == pytest.approx((a[nz] * t).sum() ** 2, rel=1e-12)
    assert vopt < PR.variance_factor(S2, None, BOX, ALPHA, Mg)


def test_training_set_fold():
    Mg = 5
    m = PR.lattice_grid(Mg)
    k = PR.kvecs(m, BOX)
    k2 = (k * k).sum(1)
    S2 = np.where(k2 > 0, 2.0 + np.abs(m[:, 0]), 0.0)  # depends on |m_x| only
    feat, y, w, mo = PR.training_set(S2, BOX, ALPHA, Mg, 1.0)
    assert len(mo) == (Mg + 1) ** 3 - 1
    assert w.sum() == pytest.approx(1.0, rel=1e-6)
    ko = PR.kvecs(mo, BOX)
    kn = np.sqrt((ko * ko).sum(1))
    t = np.sqrt(2.0 + mo[:, 0]) / kn
    yy = np.log(t)
    ww = PR.analytic_weight(mo, BOX, ALPHA) * t * (2.0 ** (mo > 0).sum(1))
    ww /= ww.sum()
    assert np.allclose(y, yy - (ww * yy).sum(), atol=2e-6)
    assert np.allclose(w, ww, rtol=2e-6)
