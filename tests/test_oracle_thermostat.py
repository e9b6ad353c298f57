"""Pins of the oracle's NVT thermostat pieces (orc_temperature, orc_scale_velocities,
orc_rescale_coeff, orc_berendsen_coeff, orc_nose_hoover_kick, orc_nose_hoover_xi,
orc_langevin_force) by closed forms: the temperature of a hand-built velocity set, the
rescale hitting T0 exactly, Berendsen's first-order relaxation T' = T + dt/tau (T0 - T),
the Nose-Hoover half step reducing to the plain half kick at xi = 0 and to a pure
friction factor at F = 0, the friction increment 5 (T/T0 - 1) per step, and the Langevin
force's two terms separately."""
import math

import numpy as np


def _rand_v(n, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.normal(0, 1.3, n) for _ in range(3)], rng.uniform(0.5, 2.5, n)


def test_temperature_closed_form(orc):
    n = 40
    v = [np.full(n, 0.7), np.zeros(n), np.full(n, -0.2)]
    m = np.full(n, 1.5)
    # sum m |v|^2 = n * 1.5 * (0.49 + 0.04); T = sum / dof (kB = 1)
    want = n * 1.5 * 0.53 / (3 * n - 3)
    assert math.isclose(orc.temperature(v, m), want, rel_tol=1e-14)
    # explicit dof (in-built lattice: 3n), kB and the mvv2e unit factor
    assert math.isclose(orc.temperature(v, m, dof=3 * n, kB=2.0, mvv2e=3.0),
                        3.0 * n * 1.5 * 0.53 / (3 * n * 2.0), rel_tol=1e-14)
    # mass NULL means unit masses
    assert math.isclose(orc.temperature(v), n * 0.53 / (3 * n - 3), rel_tol=1e-14)


def test_rescale_hits_target(orc):
    v, m = _rand_v(300)
    T = orc.temperature(v, m)
    for T0 in (0.3, 1.0, 2.7):
        c = orc.rescale_coeff(T, T0)
        assert math.isclose(c * c * T, T0, rel_tol=1e-13)
        assert math.isclose(orc.temperature(orc.scale_velocities(v, c), m), T0, rel_tol=1e-13)


def test_berendsen_first_order_relaxation(orc):
    v, m = _rand_v(257, seed=5)
    T = orc.temperature(v, m)
    T0, dt = 1.0, 0.002
    for tdamp in (0.5, 0.01, dt):
        c = orc.berendsen_coeff(T, T0, dt, tdamp)
        Tn = orc.temperature(orc.scale_velocities(v, c), m)
        assert math.isclose(Tn, T + dt / tdamp * (T0 - T), rel_tol=1e-13)
    # tdamp = dt is the plain rescale; at T = T0 the coefficient is 1
    assert math.isclose(orc.berendsen_coeff(T, T0, dt, dt), orc.rescale_coeff(T, T0),
                        rel_tol=1e-15)
    assert orc.berendsen_coeff(T0, T0, dt, 0.5) == 1.0


def test_nose_hoover_half_step(orc):
    v, m = _rand_v(100, seed=7)
    rng = np.random.default_rng(8)
    f = [rng.normal(0, 4, 100) for _ in range(3)]
    dt = 0.004
    # xi = 0: the plain half kick
    a = orc.nose_hoover_kick(v, f, dt, 0.0, m)
    b = orc.kick(v, f, dt, m)
    for d in range(3):
        assert np.allclose(a[d], b[d], rtol=1e-15, atol=0)
    # F = 0: v (1 - dt/2 xi)
    xi = 3.5
    z = [np.zeros(100)] * 3
    a = orc.nose_hoover_kick(v, z, dt, xi, m)
    for d in range(3):
        assert np.allclose(a[d], v[d] * (1 - 0.5 * dt * xi), rtol=1e-14, atol=0)
    # linear in F and v: both terms at once
    a = orc.nose_hoover_kick(v, f, dt, xi, m)
    for d in range(3):
        assert np.allclose(a[d], v[d] + 0.5 * dt * (f[d] / m - xi * v[d]), rtol=1e-14,
                           atol=1e-15)


def test_nose_hoover_friction_update(orc):
    # tauT = dt / 10, so the increment is 5 (T/T0 - 1) whatever dt is
    for dt in (0.001, 0.01):
        assert math.isclose(orc.nose_hoover_xi(0.25, 1.2, 1.0, dt), 0.25 + 5 * 0.2,
                            rel_tol=1e-14)
        assert orc.nose_hoover_xi(-0.5, 2.0, 2.0, dt) == -0.5
    assert orc.nose_hoover_xi(0.0, 0.5, 1.0, 0.002) < 0  # cold system: negative friction


def test_langevin_force_terms(orc):
    v, m = _rand_v(64, seed=9)
    rng = np.random.default_rng(10)
    f = [rng.normal(0, 2, 64) for _ in range(3)]
    gamma, kbt, dt = 100.0, 1.0, 0.002
    zero3 = np.zeros(3)
    # no noise: pure friction -m gamma v
    a = orc.langevin_force(f, v, zero3, gamma, kbt, dt, m)
    for d in range(3):
        assert np.allclose(a[d], f[d] - m * gamma * v[d], rtol=1e-14, atol=1e-12)
    # at rest: the same noise vector for every atom, amplitude sqrt(2 m gamma kBT / dt)
    g = np.array([0.3, -1.1, 2.0])
    z = [np.zeros(64)] * 3
    a = orc.langevin_force(f, z, g, gamma, kbt, dt, m)
    amp = np.sqrt(2 * m * gamma * kbt / dt)
    for d in range(3):
        assert np.allclose(a[d] - f[d], amp * g[d], rtol=1e-14, atol=0)
