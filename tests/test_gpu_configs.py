"""GPU parity at the benchmark configurations' own parameters and sizes.

Each test runs the HIP path through the C ABI on one of the workloads the
benchmark is quoted on and compares it with the CPU oracle on identical inputs:

* 100,000 ions with a learned k-space proposal (P = 50): the importance weights
  w = a(m) / (P q(m)) of the proposal's draws (host draws, bit-identical to the
  device sampler) feed the oracle's k-space force (oracle/refcpu.c
  orc_kspace_forces, the reference's ComputeNewRBEForceWorklet,
  rbmd/framework/run/worklet/RunWorklet.cpp:1508-1548, with w in place of S/P);
* the large-system physics (alpha 0.3, real-space cutoff 6.0, P = 500) at
  32,768 ions in FP32 and FP64;
* 64 uneven replicas (N_r = floor(2048 * 32^u_r)) advanced in one batch with
  learned proposals: the smallest, a middle and the largest replica against the
  oracle's velocity-Verlet loop fed the same per-step draws.

Bars as in test_gpu_parity: forces within 1e-5 relative (L2) in FP32 and 1e-12
in FP64, per-atom errors within 50x that with a floor of 1e-3 RMS(F).
"""
import math
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


def _force(Lb, n, kspace, P, alpha=0.5, rc_coul=5.0, precision=32):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    return LJCutCoulKspaceForce(ForceConfig(
        box_len=tuple(float(v) for v in Lb), cut_off=2.5, coulomb_cut=rc_coul, skin=0.3,
        alpha=alpha, kspace=kspace, coulomb_sample_num=P, precision=precision, n_max=n))


def _compute(f, x, y, z):
    X, Y, Z = _dev(x), _dev(y), _dev(z)
    F = [torch.zeros(x.size, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F, energy_virial=True)
    torch.cuda.synchronize()
    return np.stack([t.double().cpu().numpy() for t in F], 1)


def _check(F, Fo, tol, what):
    err = np.linalg.norm(F - Fo) / np.linalg.norm(Fo)
    floor = 1e-3 * np.sqrt(np.mean(np.sum(Fo ** 2, 1)))
    per = np.linalg.norm(F - Fo, axis=1) / np.maximum(np.linalg.norm(Fo, axis=1), floor)
    _log(f"  {what}: rel L2 {err:.3e}, per-atom max {per.max():.3e}")
    assert err < tol, (what, err)
    assert per.max() < 50 * tol, (what, per.max())


def _trained_proposal(Lb, alpha, n, Mg, iters=300):
    """A proposal network trained by the library's own kernels on the spectrum of
    uncorrelated charges (|rho(k)|^2 = n), tabulated on the lattice box |m_d| <= Mg."""
    from copy_gpu_amd import proposal as PR

    m = PR.lattice_grid(Mg)
    S2 = np.full(len(m), float(n))
    S2[(m == 0).all(1)] = 0.0
    prop, _, hist = PR.learned_proposal(S2, Lb, alpha, Mg, lam=0.9,
                                        schedule=((1e-2, iters),))
    assert hist[-1] < hist[0]
    return prop


# ----------------------------------------------------------------------------- learned, 100k

def test_learned_proposal_100k_ions_forces_match_oracle(orc, gpu):
    from copy_gpu_amd import systems

    n, P, alpha = 100000, 50, 0.5
    _log(f"case: learned proposal, {n} ions, P={P}, FP32 forces and energies vs oracle")
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=SEED)
    Lb = np.array([L] * 3)
    assert abs(L - 69.336) < 1e-3
    prop = _trained_proposal(Lb, alpha, n, Mg=32)
    f = _force(Lb, n, "RBE", P)
    f.Init(_dev(q))
    prop.attach(f)
    # the device draw equals the host draw, then the host draw is fed to both sides
    f.sample_kvectors(P, SEED, 17)
    md, wd, _, _ = f.kspace()
    m, w = prop.sample_host(P, SEED, 17)
    assert np.array_equal(md, m) and np.array_equal(wd, w)
    S, _ = orc.rbe_S(alpha, Lb)
    assert not np.allclose(w, S / P), "learned weights must differ from the analytic S/P"
    f.set_kvectors(_dev(m.reshape(-1), torch.int32), _dev(w))
    F = _compute(f, x, y, z)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    reo, imo = orc.rho(xw, yw, zw, q, Lb, m)
    _, _, re, im = f.kspace()
    scale = np.max(np.hypot(reo, imo))
    rerr = max(np.abs(re - reo).max(), np.abs(im - imo).max()) / scale
    _log(f"  rho max error / max|rho| = {rerr:.3e}")
    assert rerr < 1e-5
    Fp, evp = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, alpha)
    Fk, evk = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, m, w, reo, imo)
    _check(F, Fp + Fk, 1e-5, "total force")
    _check(F - Fp, Fk, 1e-4, "k-space part")
    ev = f.energy_virial()
    assert abs(ev["e_kspace"] - evk[0]) < 1e-4 * abs(evk[0])
    Wo = evp[2:8] + evk[1:7]
    assert np.linalg.norm(ev["virial"] - Wo) < 1e-4 * np.linalg.norm(Wo)
    f.close()
    prop.close()


# ----------------------------------------------------------------------------- c3 physics

@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_large_system_parameters_alpha03_rc6_p500(orc, gpu, precision, tol):
    """The 1,048,576-ion workload's physics (alpha 0.3, rc_coul 6.0, P = 500) at
    32,768 ions: the FP32 erfc at z = sqrt(0.3) 6 = 3.29 and the longer rows."""
    from copy_gpu_amd import systems

    n, P, alpha = 32768, 500, 0.3
    _log(f"case: alpha {alpha}, rc_coul 6.0, P={P}, {n} ions, FP{precision}")
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=SEED + 3)
    Lb = np.array([L] * 3)
    m, S = systems.sample_kvectors_host(alpha, Lb, P, seed=SEED, step=9)
    f = _force(Lb, n, "RBE", P, alpha=alpha, rc_coul=6.0, precision=precision)
    f.Init(_dev(q))
    f.set_kvectors(_dev(m.reshape(-1), torch.int32))
    F = _compute(f, x, y, z)
    assert f.stats()["overflow_rows"] == 0
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fp, evp = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 6.0, alpha)
    Fk, evk = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, m, np.full(P, S / P))
    _check(F, Fp + Fk, tol, "total force")
    ev = f.energy_virial()
    assert abs(ev["e_real"] - evp[1]) < 10 * tol * abs(evp[1])
    assert abs(ev["e_kspace"] - evk[0]) < 10 * tol * abs(evk[0])
    # the real-space part alone (no k-space) at the same bar
    g = _force(Lb, n, "NONE", 0, alpha=alpha, rc_coul=6.0, precision=precision)
    g.Init(_dev(q))
    _check(_compute(g, x, y, z), Fp, tol, "real-space force")
    f.close()
    g.close()


# ----------------------------------------------------------------------------- 64 replicas

def test_64_uneven_replicas_match_oracle_loop(orc, gpu):
    """64 replicas with N_r = floor(2048 * 32^u_r), each with its own learned
    proposal (P = 100), 8 batched steps in one graph; the smallest, a middle and the
    largest replica are checked against the oracle's velocity-Verlet loop fed the
    proposal's host draws of the same (seed, step)."""
    from copy_gpu_amd import ForceConfig, systems
    from copy_gpu_amd import proposal as PR
    from copy_gpu_amd.batch import ReplicaBatch

    sizes = systems.replica_sizes(64, 2048, 65536, seed=SEED)
    P, alpha, dt, steps = 100, 0.5, 0.002, 8
    _log(f"case: 64 replicas, {sum(sizes)} ions ({min(sizes)}..{max(sizes)}), learned P={P}, "
         f"{steps} steps")
    order = sorted(range(64), key=lambda r: sizes[r])
    pick = [order[0], order[32], order[-1]]
    # one network for every replica (its features are physical k), tabulated per box
    L0 = (sizes[order[-1]] / 0.3) ** (1.0 / 3.0)
    mlp, _, k_scale = PR.train_proposal(
        np.where((PR.lattice_grid(16) == 0).all(1), 0.0, float(sizes[order[-1]])),
        (L0, L0, L0), alpha, 16, schedule=((1e-2, 300),))
    syss, cfgs = [], []
    for r, n in enumerate(sizes):
        x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=SEED + 1000 + r)
        vx, vy, vz = systems.velocities(n, 1.0, seed=SEED + 1000 + r)
        syss.append((x, y, z, q, L, vx, vy, vz))
        cfgs.append(ForceConfig(box_len=(L, L, L), cut_off=2.5, coulomb_cut=5.0, skin=0.3,
                                alpha=alpha, kspace="RBE", coulomb_sample_num=P, precision=32,
                                n_max=n))
    b = ReplicaBatch(cfgs)
    props, states = [], []
    seeds = [SEED + r for r in range(64)]
    for r, ((x, y, z, q, L, vx, vy, vz), fr) in enumerate(zip(syss, b.replicas)):
        fr.Init(_dev(q))
        Mr = max(8, int(math.ceil(3.0 * math.sqrt(alpha / 2.0) * L / math.pi)))
        pr = PR.Proposal(Mr, PR.grid_density(mlp, (L, L, L), alpha, Mr, k_scale), 0.9, alpha,
                         (L, L, L))
        pr.attach(fr)
        props.append(pr)
        st = [_dev(a) for a in (x, y, z, vx, vy, vz)] + \
            [torch.zeros(x.size, device="cuda", dtype=torch.float32) for _ in range(3)] + \
            [torch.zeros(3 * x.size, device="cuda", dtype=torch.int32)]
        fr.sample_kvectors(P, seeds[r], 0)
        fr.Execute(*st[:3], *st[6:9])
        states.append(st)
    b.md_run([tuple(s) for s in states], dt=dt, nsteps=steps, seeds=seeds, step0=1,
             use_graph=True)
    torch.cuda.synchronize()
    for r in pick:
        x, y, z, q, L, vx, vy, vz = syss[r]
        Lb = np.array([L] * 3)
        lo = np.zeros(3)
        _log(f"  replica {r}: {x.size} ions, L = {L:.3f}")

        def forces(p, s):
            mk, wk = props[r].sample_host(P, seeds[r], s)
            Fp, _ = orc.pair_forces(*p, q, lo, Lb, 2.5, 5.0, alpha)
            return Fp + orc.kspace_forces(*p, q, Lb, alpha, mk, wk)[0]

        p = list(orc.wrap(x, y, z, lo, Lb))
        img = np.zeros(3 * x.size, np.int32)
        v = [vx, vy, vz]
        Fo = forces(p, 0)
        for s in range(1, steps + 1):
            v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
            p, img = orc.drift(p, v, dt, lo, Lb, img)
            Fo = forces(p, s)
            v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
        st = states[r]
        Pg = np.stack([t.cpu().numpy() for t in st[:3]], 1)
        Vg = np.stack([t.cpu().numpy() for t in st[3:6]], 1)
        Fg = np.stack([t.double().cpu().numpy() for t in st[6:9]], 1)
        Po, Vo = np.stack(p, 1), np.stack(v, 1)
        d = Pg - Po
        dpos = np.abs(d - Lb * np.rint(d / Lb)).max()
        ferr = np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo)
        _log(f"    max |dx| {dpos:.2e}, max |dv| {np.abs(Vg - Vo).max():.2e}, "
             f"force rel L2 {ferr:.2e}")
        # FP32 forces (1e-5 relative) integrated over 8 steps of dt = 0.002
        assert dpos < 1e-6
        assert np.abs(Vg - Vo).max() < 1e-3
        assert np.array_equal(st[9].cpu().numpy(), img)
        assert ferr < 1e-4
    b.close()
    for pr in props:
        pr.close()
