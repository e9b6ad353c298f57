"""Pair-force and energy/virial oracle against closed forms and derivatives."""
import math

import numpy as np

LO = np.zeros(3)


def _pair(orc, r, qi, qj, rc_lj=2.5, rc_coul=5.0, alpha=0.5, direction=(0.3, -0.5, 0.81)):
    u = np.array(direction) / np.linalg.norm(direction)
    Lb = np.array([30.0, 30.0, 30.0])
    p0 = np.array([10.0, 11.0, 12.0])
    p1 = p0 + r * u
    x, y, z = np.array([p0[0], p1[0]]), np.array([p0[1], p1[1]]), np.array([p0[2], p1[2]])
    F, ev = orc.pair_forces(x, y, z, np.array([qi, qj]), LO, Lb, rc_lj, rc_coul, alpha)
    return F, ev, u


def test_lj_single_pair_closed_form(orc):
    for r in [0.95, 1.1224620483, 1.5, 2.4]:
        F, ev, u = _pair(orc, r, 0.0, 0.0)
        mag = 24 * (2 * r ** -13 - r ** -7)  # -dU/dr, U = 4 (r^-12 - r^-6)
        np.testing.assert_allclose(F[0], -mag * u, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(F[1], mag * u, rtol=1e-13, atol=1e-13)
        assert abs(ev[0] - 4 * (r ** -12 - r ** -6)) < 1e-13
    F, ev, _ = _pair(orc, 2.6, 0.0, 0.0)  # beyond rc_lj
    assert np.all(F == 0) and ev[0] == 0


def test_coulomb_force_is_minus_energy_derivative(orc):
    h = 1e-6
    for r in [0.9, 1.7, 3.3, 4.9]:
        F, ev, u = _pair(orc, r, 1.0, -1.0, rc_lj=0.5)
        _, evp, _ = _pair(orc, r + h, 1.0, -1.0, rc_lj=0.5)
        _, evm, _ = _pair(orc, r - h, 1.0, -1.0, rc_lj=0.5)
        dE = (evp[1] - evm[1]) / (2 * h)
        # force on atom 1 along u is -dE/dr
        assert abs(F[1] @ u + dE) < 1e-7 * max(1.0, abs(dE)), (r, F[1] @ u, -dE)
        e_closed = -math.erfc(math.sqrt(0.5) * r) / r
        assert abs(ev[1] - e_closed) < 1e-14


def _random_system(n, Lb, seed, dmin=0.85):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, Lb)
        ok = True
        for q in pts:
            d = p - q
            d -= Lb * np.rint(d / Lb)
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    p = np.array(pts)
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    rng.shuffle(q)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q


def test_total_pair_force_is_gradient_and_sums_to_zero(orc):
    Lb = np.array([9.0, 10.0, 11.0])
    x, y, z, q = _random_system(60, Lb, 3)
    F, ev = orc.pair_forces(x, y, z, q, LO, Lb, 2.5, 4.0, 0.6)
    assert np.abs(F.sum(0)).max() < 1e-10
    h = 1e-6
    for i in [0, 17, 42]:
        for d, arr in enumerate((x, y, z)):
            a = arr.copy()
            a[i] += h
            args = [x, y, z]
            args[d] = a
            _, evp = orc.pair_forces(*args, q, LO, Lb, 2.5, 4.0, 0.6)
            a[i] -= 2 * h
            _, evm = orc.pair_forces(*args, q, LO, Lb, 2.5, 4.0, 0.6)
            g = ((evp[0] + evp[1]) - (evm[0] + evm[1])) / (2 * h)
            assert abs(F[i, d] + g) < 1e-6 * max(1.0, abs(g)), (i, d, F[i, d], -g)


def test_fcc_lattice_zero_force_and_shell_energy(orc):
    from copy_gpu_amd import systems

    x, y, z, Lb = systems.fcc(6, 0.8442)
    Lv = np.array([Lb] * 3)
    n = x.size
    F, ev = orc.pair_forces(x, y, z, np.zeros(n), LO, Lv, 2.5, 0.0, 0.5)
    assert np.abs(F).max() < 1e-10
    a = (4 / 0.8442) ** (1 / 3)
    shells = [(12, a / math.sqrt(2)), (6, a), (24, a * math.sqrt(1.5)), (12, a * math.sqrt(2)),
              (24, a * math.sqrt(2.5))]
    u = sum(c * 4 * (r ** -12 - r ** -6) for c, r in shells if r < 2.5)
    assert abs(ev[0] / n - 0.5 * u) < 1e-12


def test_virial_diagonal_is_strain_derivative(orc):
    """W_aa = -dE/d(eps_aa) under a homogeneous stretch of positions and box (LJ +
    real-space + Ewald k-space), and the off-diagonal pair virial under a shear of
    an isolated cluster."""
    Lb = np.array([9.0, 9.5, 10.0])
    x, y, z, q = _random_system(40, Lb, 5)
    alpha, rc = 0.9, 4.4

    def energy_and_virial(x, y, z, Lb):
        _, ev = orc.pair_forces(x, y, z, q, LO, Lb, 2.5, rc, alpha)
        m, w = orc.ewald_kvectors(Lb, alpha, [14, 14, 14])
        _, evk = orc.kspace_forces(x, y, z, q, Lb, alpha, m, w)
        e = ev[0] + ev[1] + evk[0]
        W = np.array(ev[2:8]) + np.array(evk[1:7])
        return e, W

    _, W = energy_and_virial(x, y, z, Lb)
    h = 1e-6
    for d in range(3):
        s = np.ones(3)
        s[d] += h
        ep, _ = energy_and_virial(x * s[0], y * s[1], z * s[2], Lb * s)
        s[d] -= 2 * h
        em, _ = energy_and_virial(x * s[0], y * s[1], z * s[2], Lb * s)
        g = (ep - em) / (2 * h)
        assert abs(W[d] + g) < 1e-6 * max(1.0, abs(g)), (d, W[d], -g)
    # shear of a cluster far from any image: W_xy = -dE/d(gamma), x -> x + gamma y
    Lbig = np.array([60.0, 60.0, 60.0])
    cx, cy, cz, cq = _random_system(30, np.array([6.0, 6.0, 6.0]), 9)
    cx, cy, cz = cx + 20, cy + 20, cz + 20
    _, ev = orc.pair_forces(cx, cy, cz, cq, LO, Lbig, 2.5, 4.0, 0.5)
    g = []
    for sgn in (+1, -1):
        _, e2 = orc.pair_forces(cx + sgn * h * (cy - 20), cy, cz, cq, LO, Lbig, 2.5, 4.0, 0.5)
        g.append(e2[0] + e2[1])
    dE = (g[0] - g[1]) / (2 * h)
    assert abs(ev[5] + dE) < 1e-6 * max(1.0, abs(dE))


def test_coulomb_virial_trace_equals_energy(orc):
    """Converged Ewald: E_coul is homogeneous of degree -1 in (positions, box),
    so tr(W_real + W_k) = E_real + E_k + E_self."""
    Lb = np.array([8.0, 8.0, 8.0])
    x, y, z, q = _random_system(48, Lb, 11)
    alpha = 1.3
    _, ev = orc.pair_forces(x, y, z, q, LO, Lb, 1e-3, 3.99, alpha)
    m, w = orc.ewald_kvectors(Lb, alpha, [18, 18, 18])
    _, evk = orc.kspace_forces(x, y, z, q, Lb, alpha, m, w)
    e = ev[1] + evk[0] + orc.self_energy(q, alpha)
    tr = sum(ev[2:5]) + sum(evk[1:4])
    assert abs(tr - e) < 1e-6 * abs(e), (tr, e)


def test_ewald_alpha_invariance(orc):
    """Total Coulomb energy and force do not depend on the splitting parameter."""
    Lb = np.array([8.0, 8.0, 8.0])
    x, y, z, q = _random_system(48, Lb, 12)
    res = []
    for alpha in (1.2, 1.8):
        F, ev = orc.pair_forces(x, y, z, q, LO, Lb, 1e-3, 3.99, alpha)
        m, w = orc.ewald_kvectors(Lb, alpha, [20, 20, 20])
        Fk, evk = orc.kspace_forces(x, y, z, q, Lb, alpha, m, w)
        res.append((F + Fk, ev[1] + evk[0] + orc.self_energy(q, alpha)))
    (F1, e1), (F2, e2) = res
    assert abs(e1 - e2) < 1e-7 * abs(e1)
    assert np.linalg.norm(F1 - F2) < 1e-7 * np.linalg.norm(F1)
