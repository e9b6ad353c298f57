"""k-space oracle: structure factor, RBE unbiasedness against Ewald, sampler distribution."""
import math

import numpy as np
import pytest


def _system(n, Lb, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(0, Lb[d], n) for d in range(3))
    q = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    rng.shuffle(q)
    return x, y, z, q


def test_rho_matches_fsum(orc):
    """rho(k) = sum_j q_j exp(i k.r_j) against an independent math.fsum evaluation."""
    Lb = np.array([13.0, 17.0, 19.0])
    x, y, z, q = _system(700, Lb, 4)
    m = np.array([[1, 0, 0], [0, -3, 2], [7, 5, -11], [-20, 13, 9]], np.int32)
    re, im = orc.rho(x, y, z, q, Lb, m)
    for l in range(m.shape[0]):
        k = 2 * np.pi * m[l] / Lb
        ph = [k[0] * a + k[1] * b + k[2] * c for a, b, c in zip(x, y, z)]
        rr = math.fsum(qq * math.cos(p) for qq, p in zip(q, ph))
        ii = math.fsum(qq * math.sin(p) for qq, p in zip(q, ph))
        assert abs(re[l] - rr) < 1e-11 and abs(im[l] - ii) < 1e-11


def test_rbe_force_unbiased_against_ewald(orc):
    """Mean of the RBE k-space force over independent seeded k-lists (weights S/P)
    equals the Ewald k-space force within the sampling error."""
    from copy_gpu_amd import systems

    Lb = np.array([7.0, 7.0, 7.0])
    x, y, z, q = _system(40, Lb, 5)
    alpha = 0.9
    m_all, w_all = orc.ewald_kvectors(Lb, alpha, [12, 12, 12])
    F_ew, ev_ew = orc.kspace_forces(x, y, z, q, Lb, alpha, m_all, w_all)
    P, draws = 64, 600
    acc = []
    accE = []
    for s in range(draws):
        m, S = systems.sample_kvectors_host(alpha, Lb, P, seed=77, step=s)
        F, ev = orc.kspace_forces(x, y, z, q, Lb, alpha, m, np.full(P, S / P))
        acc.append(F)
        accE.append(ev[0])
    A = np.array(acc)
    mean = A.mean(0)
    se = A.std(0, ddof=1) / math.sqrt(draws)
    z_scores = (mean - F_ew) / np.maximum(se, 1e-300)
    assert np.abs(z_scores).max() < 5.0, np.abs(z_scores).max()
    # energy estimator too
    e_mean, e_se = np.mean(accE), np.std(accE, ddof=1) / math.sqrt(draws)
    assert abs(e_mean - ev_ew[0]) < 5 * e_se
    # and the estimator is not trivially exact: a single draw differs
    assert np.linalg.norm(A[0] - F_ew) > 1e-3 * np.linalg.norm(F_ew)


def _axis_pmf(alpha, Lb, d, M=200):
    m = np.arange(-M, M + 1)
    g = [np.exp(-(math.pi ** 2) * m ** 2 / (alpha * Lb[k] ** 2)) for k in range(3)]
    G = [gg.sum() for gg in g]
    S = G[0] * G[1] * G[2] - 1.0
    others = G[(d + 1) % 3] * G[(d + 2) % 3]
    pmf = g[d] * others
    pmf[M] -= 1.0  # exclude m = 0 vector
    return m, pmf / S


@pytest.mark.parametrize("alpha,Lb", [(0.5, (47.802, 47.802, 47.802)), (0.3, (12.0, 20.0, 9.0))])
def test_sampler_marginals_chi2(alpha, Lb):
    """Host k-sampler (the same draws as the device sampler) matches the target
    p(m) ~ exp(-(2 pi m / L)^2 / (4 alpha)), m != 0 (RBEPSample.h:87-93)."""
    from scipy import stats

    from copy_gpu_amd import systems

    Lb = np.array(Lb)
    draws = []
    for s in range(40):
        m, _ = systems.sample_kvectors_host(alpha, Lb, 2000, seed=123, step=s)
        draws.append(m)
    D = np.concatenate(draws)
    assert not np.any(np.all(D == 0, axis=1))
    for d in range(3):
        mvals, pmf = _axis_pmf(alpha, Lb, d)
        counts = np.array([(D[:, d] == v).sum() for v in mvals], float)
        exp = pmf * D.shape[0]
        keep = exp > 5
        obs = np.append(counts[keep], counts[~keep].sum())
        ex = np.append(exp[keep], exp[~keep].sum())
        chi2 = ((obs - ex) ** 2 / ex).sum()
        p = stats.chi2.sf(chi2, len(obs) - 1)
        assert p > 1e-4, (d, chi2, len(obs), p)
