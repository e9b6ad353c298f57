"""Pins of the oracle's velocity-Verlet pieces (orc_kick, orc_drift with wrap and image
flags) by closed forms: velocity Verlet is exact for a constant force, its map for a
harmonic force is a known 2x2 matrix, and wrapping with image flags preserves the
unwrapped trajectory."""
import numpy as np


def _vv(orc, x, v, force, dt, nsteps, lo, Lb, mass=None):
    """Kick-drift-force-kick with the oracle's pieces; returns unwrapped x, v."""
    img = np.zeros(3 * x[0].size, np.int32)
    f = force(x, img)
    for _ in range(nsteps):
        v = orc.kick(v, f, dt, mass)
        x, img = orc.drift(x, v, dt, lo, Lb, img)
        f = force(x, img)
        v = orc.kick(v, f, dt, mass)
    im = img.reshape(-1, 3)
    xu = [x[d] + im[:, d] * Lb[d] for d in range(3)]
    return xu, v, im


def test_constant_force_is_exact(orc):
    rng = np.random.default_rng(1)
    n, dt, steps = 50, 0.01, 400
    lo, Lb = np.array([-2.0, 0.0, 1.0]), np.array([5.0, 7.0, 3.0])
    x0 = [lo[d] + rng.uniform(0, Lb[d], n) for d in range(3)]
    v0 = [rng.normal(0, 3, n) for _ in range(3)]
    a = [rng.normal(0, 2, n) for _ in range(3)]
    mass = rng.uniform(0.5, 2.0, n)
    F = [a[d] * mass for d in range(3)]
    xu, v, im = _vv(orc, x0, v0, lambda x, i: F, dt, steps, lo, Lb, mass)
    t = dt * steps
    for d in range(3):
        assert np.allclose(xu[d], x0[d] + v0[d] * t + 0.5 * a[d] * t * t, rtol=0, atol=1e-10)
        assert np.allclose(v[d], v0[d] + a[d] * t, rtol=0, atol=1e-11)
    assert np.abs(im).max() >= 2  # trajectories crossed the box several times


def test_harmonic_map(orc):
    # velocity Verlet on x'' = -w^2 x is the linear map
    # M = [[1 - h, dt], [-w^2 dt (1 - h/2), 1 - h]], h = w^2 dt^2 / 2
    w, dt, steps = 1.3, 0.05, 300
    lo, Lb = np.zeros(3), np.array([1e3, 1e3, 1e3])
    x0 = [np.array([500.0 + 0.7]), np.array([500.0 - 0.2]), np.array([500.0])]
    v0 = [np.array([0.1]), np.array([0.5]), np.array([-0.3])]

    def force(x, img):
        return [-(w * w) * (x[d] + img.reshape(-1, 3)[:, d] * Lb[d] - 500.0) for d in range(3)]

    xu, v, _ = _vv(orc, x0, v0, force, dt, steps, lo, Lb)
    h = 0.5 * (w * dt) ** 2
    M = np.array([[1 - h, dt], [-(w * w) * dt * (1 - h / 2), 1 - h]])
    Mn = np.linalg.matrix_power(M, steps)
    for d in range(3):
        ref = Mn @ np.array([x0[d][0] - 500.0, v0[d][0]])
        assert abs(xu[d][0] - 500.0 - ref[0]) < 1e-9
        assert abs(v[d][0] - ref[1]) < 1e-9


def test_wrap_keeps_unwrapped_path(orc):
    rng = np.random.default_rng(5)
    n = 200
    lo, Lb = np.array([0.0, -1.0, 2.0]), np.array([3.0, 4.0, 5.0])
    x = [lo[d] + rng.uniform(0, Lb[d], n) for d in range(3)]
    v = [rng.normal(0, 40, n) for _ in range(3)]
    img = np.zeros(3 * n, np.int32)
    xw, img = orc.drift(x, v, 0.25, lo, Lb, img)
    im = img.reshape(-1, 3)
    for d in range(3):
        assert np.all(xw[d] >= lo[d]) and np.all(xw[d] < lo[d] + Lb[d])
        assert np.allclose(xw[d] + im[:, d] * Lb[d], x[d] + 0.25 * v[d], rtol=0, atol=1e-12)
