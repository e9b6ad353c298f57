"""GPU tests of the fused MD step (rho(k) partials + cell sort, rho(k) final + tile
pass, k-space force + kicks + drift) against the oracle's velocity-Verlet loop.

The oracle loop is kick / drift / forces / kick with FP64 pair forces and the same
k-space term (oracle.pair_forces, oracle.kspace_forces, oracle.kick, oracle.drift).
These runs use a small skin so the list is rebuilt many times, which exercises the
tile pass's binning for the next rebuild and the one-kernel cell sort; they cover
two LJ types, per-atom masses, image flags, full Ewald k-space and the
global-memory tile path.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261202


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _make(Lb, n, precision, kspace, P=100, skin=0.3, eps=(1.0,), sigma=(1.0,), tile_stage=0,
          ewald_tol=1e-12, alpha=0.5):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    return LJCutCoulKspaceForce(ForceConfig(
        box_len=tuple(float(v) for v in Lb), cut_off=2.5, coulomb_cut=5.0, skin=skin,
        alpha=alpha, kspace=kspace, coulomb_sample_num=P, precision=precision, eps=eps,
        sigma=sigma, n_max=n, ewald_tol=ewald_tol, tile_stage=tile_stage))


def _halfspace_lattice(Lb, alpha, tol):
    # the handle's Ewald k-list: half space m_x > 0, or m_x = 0 and (m_y, m_z) > 0,
    # k^2 <= -4 alpha ln(tol), weight 2 exp(-k^2 / 4 alpha)
    kmax2 = -4 * alpha * np.log(tol)
    M = [int(np.floor(np.sqrt(kmax2) * Lb[d] / (2 * np.pi))) + 1 for d in range(3)]
    a, b, g = np.meshgrid(np.arange(0, M[0] + 1), np.arange(-M[1], M[1] + 1),
                          np.arange(-M[2], M[2] + 1), indexing="ij")
    a, b, g = a.ravel(), b.ravel(), g.ravel()
    keep = ~((a == 0) & ((b < 0) | ((b == 0) & (g <= 0))))
    a, b, g = a[keep], b[keep], g[keep]
    k2 = ((2 * np.pi * a / Lb[0]) ** 2 + (2 * np.pi * b / Lb[1]) ** 2
          + (2 * np.pi * g / Lb[2]) ** 2)
    keep = k2 <= kmax2
    return (np.stack([a[keep], b[keep], g[keep]], 1).astype(np.int32),
            2 * np.exp(-k2[keep] / (4 * alpha)))


def _oracle_run(orc, x, y, z, q, vx, vy, vz, Lb, alpha, steps, dt, kforces, types=None,
                eps=(1.0,), sigma=(1.0,), mass=None):
    lo = np.zeros(3)

    def forces(p, s):
        Fp, _ = orc.pair_forces(*p, q, lo, Lb, 2.5, 5.0, alpha, types=types, eps=eps,
                                sigma=sigma)
        return Fp + kforces(p, s)

    p = list(orc.wrap(x, y, z, lo, Lb))
    img = np.zeros(3 * x.size, np.int32)
    v = [vx, vy, vz]
    Fo = forces(p, 0)
    for s in range(1, steps + 1):
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt, mass)
        p, img = orc.drift(p, v, dt, lo, Lb, img)
        Fo = forces(p, s)
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt, mass)
    return np.stack(p, 1), np.stack(v, 1), img, Fo


def _gpu_run(f, x, y, z, q, vx, vy, vz, steps, dt, types=None, mass=None, P=None,
             use_graph=True):
    n = x.size
    f.Init(_dev(q), None if types is None else _dev(types, torch.int32))
    if P is not None:
        f.sample_kvectors(P, SEED, 0)
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    img = torch.zeros(3 * n, device="cuda", dtype=torch.int32)
    M = None if mass is None else _dev(mass)
    b0 = f.stats()["builds"]
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1, img=img, mass=M,
             use_graph=use_graph)
    torch.cuda.synchronize()
    builds = f.stats()["builds"] - b0
    P_ = np.stack([t.cpu().numpy() for t in (X, Y, Z)], 1)
    V_ = np.stack([t.cpu().numpy() for t in (VX, VY, VZ)], 1)
    F_ = np.stack([t.double().cpu().numpy() for t in F], 1)
    return P_, V_, img.cpu().numpy(), F_, builds


def _dpos(a, b, Lb):
    d = a - b
    return np.abs(d - Lb * np.rint(d / Lb)).max()


def test_md_rebuilds_two_types_masses_ewald_fp64(orc, gpu):
    """FP64, full Ewald k-space, two LJ types, per-atom masses, skin 0.15 (a list
    rebuild nearly every step: the sigma-1.3 contacts of the random start move atoms
    up to 0.062 per step, under half the skin): 40 graph-replayed steps against the
    oracle loop."""
    from copy_gpu_amd import systems

    _log("case: FP64 Ewald MD, two types, masses, skin 0.15, 40 steps, 2000 ions")
    n, dt, steps, alpha, tol = 2000, 0.002, 40, 0.5, 1e-12
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=51)
    vx, vy, vz = systems.velocities(n, 1.0, seed=51)
    Lb = np.array([L] * 3)
    types = (np.arange(n) % 3 == 0).astype(np.int32)
    eps, sig = (1.0, 0.6), (1.0, 1.3)
    mass = 1.0 + 0.5 * (np.arange(n) % 2)
    f = _make(Lb, n, 64, "EWALD", skin=0.15, eps=eps, sigma=sig, ewald_tol=tol, alpha=alpha)
    Pg, Vg, Ig, Fg, builds = _gpu_run(f, x, y, z, q, vx, vy, vz, steps, dt, types=types,
                                      mass=mass)
    _log(f"  list builds during the run: {builds}")
    assert builds >= 5
    m, w = _halfspace_lattice(Lb, alpha, tol)
    assert f.stats()["ewald_k"] == m.shape[0]
    Po, Vo, Io, Fo = _oracle_run(
        orc, x, y, z, q, vx, vy, vz, Lb, alpha, steps, dt,
        lambda p, s: orc.kspace_forces(*p, q, Lb, alpha, m, w)[0], types=types, eps=eps,
        sigma=sig, mass=mass)
    assert _dpos(Pg, Po, Lb) < 1e-9
    assert np.abs(Vg - Vo).max() < 1e-8
    assert np.array_equal(Ig, Io)
    assert np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo) < 1e-9


def test_md_rbe_fp32_rebuilds(orc, gpu):
    """FP32 forces with device-sampled RBE k-lists (one per step), skin 0.1: 30
    steps against the oracle loop fed the host draws of the same (seed, step)."""
    from copy_gpu_amd import systems

    _log("case: FP32 RBE MD, skin 0.1, 30 steps, 3000 ions")
    n, dt, steps, alpha, P = 3000, 0.002, 30, 0.5, 100
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=52)
    vx, vy, vz = systems.velocities(n, 1.0, seed=52)
    Lb = np.array([L] * 3)
    f = _make(Lb, n, 32, "RBE", P=P, skin=0.1)
    Pg, Vg, Ig, Fg, builds = _gpu_run(f, x, y, z, q, vx, vy, vz, steps, dt, P=P)
    _log(f"  list builds during the run: {builds}")
    assert builds >= 4

    def kf(p, s):
        mk, S = systems.sample_kvectors_host(alpha, Lb, P, SEED, s)
        return orc.kspace_forces(*p, q, Lb, alpha, mk, np.full(P, S / P))[0]

    Po, Vo, Io, Fo = _oracle_run(orc, x, y, z, q, vx, vy, vz, Lb, alpha, steps, dt, kf)
    # FP32 forces (1e-5 relative) integrated over 30 steps of dt = 0.002
    assert _dpos(Pg, Po, Lb) < 1e-6
    assert np.abs(Vg - Vo).max() < 1e-3
    assert np.array_equal(Ig, Io)
    assert np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo) < 1e-4


@pytest.mark.parametrize("stage", [64])
def test_md_global_memory_tiles_equal_staged(gpu, stage):
    """Every tile read from global memory (LDS stage 64 atoms) gives the same bits
    over an MD run, rebuilds included, as the LDS-staged tiles."""
    from copy_gpu_amd import systems

    _log(f"case: MD with LDS stage {stage} vs default, FP32 RBE, skin 0.1, 20 steps")
    n, dt, steps, P = 4000, 0.002, 20, 100
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=53)
    vx, vy, vz = systems.velocities(n, 1.0, seed=53)
    Lb = np.array([L] * 3)
    outs = []
    for ts in (0, stage):
        f = _make(Lb, n, 32, "RBE", P=P, skin=0.1, tile_stage=ts)
        res = _gpu_run(f, x, y, z, q, vx, vy, vz, steps, dt, P=P)
        if ts:
            assert f.stats()["big_tiles"] > 0
        outs.append(res[:4])
        f.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_md_lj_fcc_fp64_matches_oracle_and_conserves_energy(orc, gpu):
    """The reference's CPU-runnable case: 4000 LJ atoms on an FCC lattice at density
    0.8442, no charges, FP64, dt 0.005.  200 steps match the oracle loop; over 1000
    steps the total energy stays within 5e-3 of its start (velocity Verlet with an
    unshifted cutoff: bounded fluctuation, no drift)."""
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce, systems

    _log("case: LJ FCC FP64 MD, 4000 atoms, 200 steps vs oracle, 1000 steps energy")
    x, y, z, L = systems.fcc(10, 0.8442)
    n = x.size
    q = np.zeros(n)
    vx, vy, vz = systems.velocities(n, 1.0, seed=61)
    Lb = np.array([L] * 3)
    dt = 0.005
    f = LJCutCoulKspaceForce(ForceConfig(box_len=(L, L, L), cut_off=2.5, coulomb_cut=0.0,
                                         skin=0.3, alpha=0.0, kspace="NONE", precision=64,
                                         n_max=n))
    Pg, Vg, Ig, Fg, builds = _gpu_run(f, x, y, z, q, vx, vy, vz, 200, dt)
    _log(f"  list builds during the first 200 steps: {builds}")
    assert builds >= 1
    Po, Vo, Io, Fo = _oracle_run(orc, x, y, z, q, vx, vy, vz, Lb, 0.0, 200, dt,
                                 lambda p, s: np.zeros((n, 3)))
    assert _dpos(Pg, Po, Lb) < 1e-8
    assert np.abs(Vg - Vo).max() < 1e-7
    assert np.array_equal(Ig, Io)

    def energy(P_, V_):
        X, Y, Z = (_dev(P_[:, d]) for d in range(3))
        F = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
        f.Execute(X, Y, Z, *F, energy_virial=True)
        ev = f.energy_virial()
        return ev["e_lj"] + 0.5 * float((V_ ** 2).sum())

    e0 = energy(np.stack([x, y, z], 1), np.stack([vx, vy, vz], 1))
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=1000, seed=0, step0=1)
    torch.cuda.synchronize()
    P1 = np.stack([t.cpu().numpy() for t in (X, Y, Z)], 1)
    V1 = np.stack([t.cpu().numpy() for t in (VX, VY, VZ)], 1)
    e1 = energy(P1, V1)
    _log(f"  total energy {e0:.6f} -> {e1:.6f} after 1000 steps")
    assert abs(e1 - e0) < 5e-3 * abs(e0)


def _oracle_thermo_run(orc, x, y, z, q, vx, vy, vz, Lb, alpha, steps, dt, kforces, mass,
                       kind, T0, tdamp, gamma, kbt, noise):
    """The reference's NVT step order (ExecutionNVT::Solve): kick, drift, force (+ the
    Langevin term), kick, temperature, thermostat."""
    lo = np.zeros(3)

    def forces(p):
        Fp, _ = orc.pair_forces(*p, q, lo, Lb, 2.5, 5.0, alpha)
        return Fp + kforces(p)

    p = list(orc.wrap(x, y, z, lo, Lb))
    img = np.zeros(3 * x.size, np.int32)
    v = [vx, vy, vz]
    F = forces(p)
    F = [F[:, 0], F[:, 1], F[:, 2]]
    xi, Ts = 0.0, []
    for s in range(steps):
        v = orc.kick(v, F, dt, mass)
        p, img = orc.drift(p, v, dt, lo, Lb, img)
        F = forces(p)
        F = [F[:, 0], F[:, 1], F[:, 2]]
        if kind == "LANGEVIN":
            F = orc.langevin_force(F, v, noise(s), gamma, kbt, dt, mass)
        v = orc.kick(v, F, dt, mass)
        T = orc.temperature(v, mass)
        Ts.append(T)
        if kind == "BERENDSEN":
            v = orc.scale_velocities(v, orc.berendsen_coeff(T, T0, dt, tdamp))
        elif kind == "TEMP_RESCALE":
            v = orc.scale_velocities(v, orc.rescale_coeff(T, T0))
        elif kind == "NOSE_HOOVER":
            v = orc.nose_hoover_kick(v, F, dt, xi, mass)
            xi = orc.nose_hoover_xi(xi, T, T0, dt)
    return np.stack(p, 1), np.stack(v, 1), img, np.array(Ts), xi


@pytest.mark.parametrize("kind", ["BERENDSEN", "TEMP_RESCALE", "NOSE_HOOVER", "LANGEVIN"])
def test_md_thermostat_matches_oracle(orc, gpu, kind):
    """FP64 Ewald NVT run of 40 steps (a 32-step graph and an 8-step remainder
    graph) with per-atom masses and list rebuilds, against the oracle's loop in the
    reference's order: positions, velocities, image flags, the temperature of every
    step, and the Nose-Hoover friction."""
    from copy_gpu_amd import systems

    _log(f"case: NVT {kind}, FP64 Ewald, 1500 ions, 40 steps")
    n, dt, steps, alpha, tol = 1500, 0.002, 40, 0.5, 1e-12
    T0, tdamp, gamma, kbt, nseed, ns0 = 1.1, 0.05, 100.0, 1.0, 77, 5
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=61)
    vx, vy, vz = systems.velocities(n, 1.0, seed=61)
    Lb = np.array([L] * 3)
    mass = 1.0 + 0.25 * (np.arange(n) % 3)
    f = _make(Lb, n, 64, "EWALD", skin=0.1, ewald_tol=tol, alpha=alpha)
    f.Init(_dev(q))
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    img = torch.zeros(3 * n, device="cuda", dtype=torch.int32)
    log = torch.zeros(steps + 2, device="cuda", dtype=torch.float64)
    f.set_thermostat(kind, (T0, T0, tdamp), gamma=gamma, kbt=kbt, seed=nseed, noise_step=ns0)
    f.set_temperature_log(log)
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1, img=img,
             mass=_dev(mass), use_graph=True)
    torch.cuda.synchronize()
    st = f.thermostat_state()
    Pg = np.stack([t.cpu().numpy() for t in (X, Y, Z)], 1)
    Vg = np.stack([t.cpu().numpy() for t in (VX, VY, VZ)], 1)
    Tg = log.cpu().numpy()
    m, w = _halfspace_lattice(Lb, alpha, tol)
    Po, Vo, Io, To, xio = _oracle_thermo_run(
        orc, x, y, z, q, vx, vy, vz, Lb, alpha, steps, dt,
        lambda p: orc.kspace_forces(*p, q, Lb, alpha, m, w)[0], mass, kind, T0, tdamp, gamma,
        kbt, lambda s: systems.langevin_noise(nseed, ns0 + s))
    _log(f"  T first/last {To[0]:.4f}/{To[-1]:.4f}; max |dT|/T "
         f"{np.abs(Tg[:steps] - To).max() / To.max():.2e}; xi {st['xi']:.6f} vs {xio:.6f}")
    assert _dpos(Pg, Po, Lb) < 1e-9
    assert np.abs(Vg - Vo).max() < 1e-8
    assert np.array_equal(img.cpu().numpy(), Io)
    assert np.allclose(Tg[:steps], To, rtol=1e-11, atol=0)
    assert np.all(Tg[steps:] == 0)  # entries past the run's steps stay untouched
    assert st["T"] == Tg[steps - 1]
    if kind == "NOSE_HOOVER":
        assert abs(st["xi"] - xio) <= 1e-9 * max(1.0, abs(xio))
    if kind == "LANGEVIN":
        assert st["noise_step"] == ns0 + steps
    if kind == "TEMP_RESCALE":
        # the velocities returned are rescaled to T0 exactly (up to rounding)
        assert abs(orc.temperature([Vg[:, 0], Vg[:, 1], Vg[:, 2]], mass) - T0) < 1e-12


def test_md_thermostat_graph_equals_eager_and_resets(gpu):
    """A Berendsen run gives the same bits replayed from graphs and launched eagerly,
    and switching the thermostat off restores the plain step."""
    from copy_gpu_amd import systems

    _log("case: NVT Berendsen graph vs eager, then thermostat off, FP32 RBE, 3000 ions")
    n, dt, steps = 3000, 0.002, 20
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=62)
    vx, vy, vz = systems.velocities(n, 1.0, seed=62)
    Lb = np.array([L] * 3)

    def run(use_graph, thermo):
        f = _make(Lb, n, 32, "RBE", P=100, skin=0.2)
        f.Init(_dev(q))
        f.sample_kvectors(100, SEED, 0)
        X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
        F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
        f.Execute(X, Y, Z, *F)
        if thermo is not None:
            f.set_thermostat(thermo, (1.3, 1.3, 0.1))
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1,
                 use_graph=use_graph)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in (X, Y, Z, VX, VY, VZ)], f

    a, _ = run(True, "BERENDSEN")
    b, _ = run(False, "BERENDSEN")
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    c, _ = run(True, "NONE")
    d, _ = run(True, None)
    for u, v in zip(c, d):
        assert np.array_equal(u, v)
    # the thermostat acted: the velocities differ from the plain run's
    assert not np.array_equal(a[3], c[3])


def test_md_thermostat_rejects_bad_parameters(gpu):
    from copy_gpu_amd import systems
    from copy_gpu_amd.force import RBMDError

    _log("case: thermostat parameter errors")
    x, y, z, q, L = systems.random_ions(1500, 0.3, 0.8, seed=63)
    f = _make(np.array([L] * 3), 1500, 32, "NONE")
    with pytest.raises(RBMDError, match="t_damp"):
        f.set_thermostat("BERENDSEN", (1.0, 1.0, 0.0))
    with pytest.raises(RBMDError, match="t_target"):
        f.set_thermostat("TEMP_RESCALE", (0.0, 0.0, 0.5))
    with pytest.raises(RBMDError, match="dof"):
        f.set_thermostat("TEMP_RESCALE", (1.0, 1.0, 0.5), dof=-1.0)


def test_md_thermostat_large_system_reduction(gpu):
    """140,000 ions (more closing-kick workgroups than the opening kick re-sums, so the
    one-workgroup reduction runs after every step): under TEMP_RESCALE every step's
    velocities leave at T0 exactly, so the returned velocities' temperature is T0 and
    graph and eager runs agree bit for bit."""
    from copy_gpu_amd import systems

    _log("case: NVT TEMP_RESCALE, 140000 ions, per-step reduction kernel, 10 steps")
    n, dt, steps, T0 = 140000, 0.002, 10, 1.2
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=64)
    vx, vy, vz = systems.velocities(n, 1.0, seed=64)
    Lb = np.array([L] * 3)
    outs = []
    for use_graph in (True, False):
        f = _make(Lb, n, 32, "RBE", P=100, skin=0.3)
        f.Init(_dev(q))
        f.sample_kvectors(100, SEED, 0)
        X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
        F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
        f.Execute(X, Y, Z, *F)
        log = torch.zeros(steps, device="cuda", dtype=torch.float64)
        f.set_thermostat("TEMP_RESCALE", (T0, T0, 0.5))
        f.set_temperature_log(log)
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1,
                 use_graph=use_graph)
        torch.cuda.synchronize()
        outs.append(([t.cpu().numpy() for t in (X, Y, Z, VX, VY, VZ)], log.cpu().numpy()))
    (a, la), (b, lb) = outs
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(la, lb)
    v = a[3:]
    T = sum(float(np.dot(c, c)) for c in v) / (3 * n - 3)
    _log(f"  T per step {la[0]:.6f} .. {la[-1]:.6f}; returned {T:.15f}")
    assert abs(T - T0) < 1e-12
    # every step logged a temperature (the random packing heats within each step, so
    # the closing-kick values sit above T0)
    assert np.all(np.isfinite(la)) and np.all(la > 0) and np.unique(la).size == steps


def test_thermo_output_row_matches_oracle(orc, gpu):
    """The reference's thermo row after an MD run with the measure-only thermostat:
    per-atom LJ, near and far (k-space + self) energies against the oracle's FP64 Ewald
    energies at the returned positions, the kinetic energy and temperature against the
    returned velocities (untouched by the thermostat), Residual between two rows."""
    from copy_gpu_amd import systems, thermo

    _log("case: thermo rows, FP64 Ewald, 1500 ions, MEASURE thermostat")
    n, dt, alpha, tol = 1500, 0.002, 0.5, 1e-12
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=65)
    vx, vy, vz = systems.velocities(n, 1.0, seed=65)
    Lb = np.array([L] * 3)
    mass = 1.0 + 0.5 * (np.arange(n) % 2)
    f = _make(Lb, n, 64, "EWALD", skin=0.3, ewald_tol=tol, alpha=alpha)
    f.Init(_dev(q))
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    f.set_thermostat("MEASURE")
    M = _dev(mass)
    rows = []
    for k in range(2):
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=10, seed=SEED, step0=1 + 10 * k,
                 mass=M)
        rows.append(thermo.thermo_row(f, X, Y, Z, 10 * (k + 1), dt,
                                      rows[-1] if rows else None))
    torch.cuda.synchronize()
    r = rows[-1]
    p = [t.cpu().numpy() for t in (X, Y, Z)]
    v = [t.cpu().numpy() for t in (VX, VY, VZ)]
    _, evp = orc.pair_forces(*p, q, np.zeros(3), Lb, 2.5, 5.0, alpha)
    e_self = orc.self_energy(q, alpha)
    assert abs(r["VanderWaalsEnergy"] - evp[0] / n) < 1e-11 * abs(evp[0] / n)
    assert abs(r["NearCoulombicEnergy"] - evp[1] / n) < 1e-11 * abs(evp[1] / n)
    # far energy: the reference's Kmax cube (ThermoOutput.cpp:218-252), default Kmax 2,
    # every m != 0 with weight exp(-k^2 / 4 alpha) -- not the handle's own k-space
    r2 = np.arange(-2, 3)
    cube = np.stack(np.meshgrid(r2, r2, r2, indexing="ij"), -1).reshape(-1, 3)
    cube = cube[(cube != 0).any(1)].astype(np.int32)
    kk = (2 * np.pi * cube / Lb) ** 2
    _, evk = orc.kspace_forces(*p, q, Lb, alpha, cube, np.exp(-kk.sum(1) / (4 * alpha)))
    far = (evk[0] + e_self) / n
    assert abs(r["FarCoulombicEnergy"] - far) < 1e-10 * abs(far)
    # an RBE handle reports the same deterministic far energy at these positions
    g = _make(Lb, n, 32, "RBE", alpha=alpha)
    g.Init(_dev(q))
    g.sample_kvectors(100, SEED, 0)
    g.set_thermostat("MEASURE")
    rg = thermo.thermo_row(g, X, Y, Z, 20, dt)
    assert abs(rg["FarCoulombicEnergy"] - far) < 1e-10 * abs(far)
    assert abs(thermo.far_kspace_energy(f, X, Y, Z, kmax=3) / n - far + e_self / n) > 0
    ke = 0.5 * sum(float(np.dot(mass * c, c)) for c in v) / n
    assert abs(r["KinticEnergy"] - ke) < 1e-12 * ke
    assert abs(r["Temperature"] - orc.temperature(v, mass)) < 1e-12 * r["Temperature"]
    assert r["TotalEnergy"] == r["PotentialEnergy"] + r["KinticEnergy"]
    assert r["Residual"] == r["PotentialEnergy"] - rows[0]["PotentialEnergy"]
    text = thermo.format_rows(rows)
    assert text.splitlines()[0].split() == list(thermo.COLUMNS)
    assert float(text.splitlines()[2].split()[2]) == r["VanderWaalsEnergy"]
