"""The timed CPU baseline (oracle/faithful.c, the reference's per-step algorithm in
its own shape) against the FP64 oracle (oracle/refcpu.c) on identical inputs.

faithful.c rebuilds the cell list and a skin-less list at rc every step,
materialises the N*P structure-factor arrays and reduces them by k, as the
reference does (rbmd/framework/run/ExecutionMD.cpp:412-448, 670-705,
rbmd/framework/run/worklet/RunWorklet.cpp:40-96, 170-237, 1508-1548, 1616-1677).
Its FP64 build must reproduce the oracle's velocity-Verlet loop to rounding, its
FP32 build (the reference's default `Real`) to FP32 accuracy: the baseline times
work that gives the oracle's answer.
"""
import numpy as np
import pytest


def _loop(orc, x, y, z, q, v, Lb, alpha, rc_lj, rc_coul, dt, klists, S, P):
    lo = np.zeros(3)

    def forces(p, s):
        qq = np.zeros(x.size) if q is None else q
        F, _ = orc.pair_forces(*p, qq, lo, Lb, rc_lj, rc_coul, max(alpha, 1e-3))
        if P and q is not None:
            F = F + orc.kspace_forces(*p, q, Lb, alpha, klists[s], np.full(P, S / P))[0]
        return F

    p = list(orc.wrap(x, y, z, lo, Lb))
    F0 = forces(p, 0)
    F = F0
    for s in range(1, len(klists) if P else 4):
        v = orc.kick(v, [F[:, 0], F[:, 1], F[:, 2]], dt)
        p, _ = orc.drift(p, v, dt, lo, Lb)
        F = forces(p, s)
        v = orc.kick(v, [F[:, 0], F[:, 1], F[:, 2]], dt)
    return F0, np.stack(p, 1), np.stack(v, 1), F


def _faithful(orc, x, y, z, q, v, F0, Lb, alpha, rc_lj, rc_coul, dt, klists, P, dtype, steps):
    pos = np.ascontiguousarray(np.stack([x, y, z], 1), dtype)
    vel = np.ascontiguousarray(np.stack(v, 1), dtype)
    frc = np.ascontiguousarray(F0, dtype)
    m = np.concatenate(klists[1:steps + 1]) if P else None
    rc = orc.faithful_run(pos, vel, frc, q, np.zeros(3), Lb, rc_lj, rc_coul, alpha, P, m, steps,
                          dt, 4)
    assert rc == 0
    return pos.astype(np.float64), vel.astype(np.float64), frc.astype(np.float64)


def _dpos(a, b, Lb):
    d = a - b
    return np.abs(d - Lb * np.rint(d / Lb)).max()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-4)])
def test_faithful_rbe_steps_match_oracle(orc, dtype, tol):
    from copy_gpu_amd import systems

    n, P, alpha, dt, steps = 3000, 100, 0.5, 0.002, 3
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=41)
    vx, vy, vz = systems.velocities(n, 1.0, seed=41)
    Lb = np.array([L] * 3)
    draws = [systems.sample_kvectors_host(alpha, Lb, P, 7, s) for s in range(steps + 1)]
    klists = [d[0] for d in draws]
    S = draws[0][1]
    F0, Po, Vo, Fo = _loop(orc, x, y, z, q, [vx, vy, vz], Lb, alpha, 2.5, 5.0, dt, klists, S, P)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Pf, Vf, Ff = _faithful(orc, xw, yw, zw, q, [vx, vy, vz], F0, Lb, alpha, 2.5, 5.0, dt, klists,
                           P, dtype, steps)
    ferr = np.linalg.norm(Ff - Fo) / np.linalg.norm(Fo)
    assert ferr < tol, ferr
    assert _dpos(Pf, Po, Lb) < 1e3 * tol * dt
    assert np.abs(Vf - Vo).max() < 1e2 * tol * np.abs(Vo).max()


def test_faithful_fp64_lj_fcc_matches_oracle(orc):
    """The LJ-only FP64 path (no charges) on the FCC lattice of the 4000-atom LJ
    workload, three steps from Maxwell-Boltzmann velocities."""
    from copy_gpu_amd import systems

    x, y, z, L = systems.fcc(10, 0.8442)
    vx, vy, vz = systems.velocities(x.size, 1.0, seed=5)
    Lb = np.array([L] * 3)
    dt = 0.005
    F0, Po, Vo, Fo = _loop(orc, x, y, z, None, [vx, vy, vz], Lb, 0.0, 2.5, 0.0, dt, None, 0, 0)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Pf, Vf, Ff = _faithful(orc, xw, yw, zw, None, [vx, vy, vz], F0, Lb, 0.0, 2.5, 0.0, dt, None,
                           0, np.float64, 3)
    assert np.linalg.norm(Ff - Fo) / np.linalg.norm(Fo) < 1e-10
    assert _dpos(Pf, Po, Lb) < 1e-12
    assert np.abs(Vf - Vo).max() < 1e-10
