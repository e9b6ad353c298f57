"""GPU tests of the random batch list near field and of boxes narrower than three list
radii, through the C ABI, against the CPU oracle.

Random batch list (reference: ComputeRBLNeighboursOnceWorklet and
ComputeNearForceRBLWorklet, rbmd/framework/run/worklet/RunWorklet.cpp:634-708, 957-1048;
ExecutionMD::ComputeRBLNearForce, rbmd/framework/run/ExecutionMD.cpp:460-566): the
shell selection is a counter-based draw of (seed, atom ids, list build number), so the
oracle (orc_rbl_forces) computes the same sample; forces within 1e-5 (FP32) / 1e-12
(FP64) relative, energies alike.  Small boxes: the 27-cell stencil over one or two
cells per axis lists every periodic image within the list radius (orc_neighbors_images).
"""
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261202
REF_JSON = os.path.join(os.path.dirname(__file__), "golden", "lj_salt_rbmd.json")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _make(Lb, n, precision=32, kspace="NONE", P=100, rc_lj=2.5, rc_coul=5.0, alpha=0.5,
          r_core=2.5, pice=3, seed=77, rbl=True, skin=0.3):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    return LJCutCoulKspaceForce(ForceConfig(
        box_len=tuple(float(v) for v in Lb), cut_off=rc_lj, coulomb_cut=rc_coul, skin=skin,
        alpha=alpha, kspace=kspace, coulomb_sample_num=P, precision=precision, n_max=n,
        neighbor_type="RBL" if rbl else "VERLET-LIST", r_core=r_core, rbl_pice=pice,
        rbl_seed=seed))


def _compute(f, x, y, z, ev=True):
    X, Y, Z = _dev(x), _dev(y), _dev(z)
    F = [torch.zeros(x.size, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F, energy_virial=ev)
    torch.cuda.synchronize()
    return np.stack([t.double().cpu().numpy() for t in F], 1)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_rbl_forces_match_oracle(orc, gpu, precision, tol):
    from copy_gpu_amd import systems

    n, pice, seed = 8000, 3, 77
    _log(f"case: random batch list FP{precision}, {n} ions, r_core 2.5, pice {pice}")
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=SEED + 5)
    Lb = np.array([L] * 3)
    f = _make(Lb, n, precision, pice=pice, seed=seed)
    f.Init(_dev(q))
    F = _compute(f, x, y, z)
    epoch = f.stats()["builds"]
    assert epoch == 1
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, evo = orc.rbl_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5, 2.5, pice, seed,
                             epoch)
    err = _rel(F, Fo)
    _log(f"  rel L2 force error {err:.3e}")
    assert err < tol
    floor = 1e-3 * np.sqrt(np.mean(np.sum(Fo ** 2, 1)))
    per = np.linalg.norm(F - Fo, axis=1) / np.maximum(np.linalg.norm(Fo, axis=1), floor)
    assert per.max() < 50 * tol
    assert np.abs(F.sum(0)).max() < 1e-3 * np.abs(F).max() * (1 if precision == 32 else 1e-6)
    ev = f.energy_virial()
    assert abs(ev["e_lj"] - evo[0]) < 10 * tol * abs(evo[0])
    assert abs(ev["e_real"] - evo[1]) < 10 * tol * abs(evo[1])
    assert np.linalg.norm(ev["virial"] - evo[2:8]) < 10 * tol * np.linalg.norm(evo[2:8])
    # the list is shorter than the full list (the shell is thinned to 1/pice)
    g = _make(Lb, n, precision, rbl=False)
    g.Init(_dev(q))
    _compute(g, x, y, z, ev=False)
    full, thin = g.stats()["entries"], f.stats()["entries"]
    _log(f"  list entries: full {full}, random batch {thin}")
    assert thin < 0.6 * full
    # a new build draws a new shell sample: forces at the same positions change
    f.Init(_dev(q))  # set_atoms forces a rebuild at the next compute
    F2 = _compute(f, x, y, z)
    assert f.stats()["builds"] == 2
    Fo2, _ = orc.rbl_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5, 2.5, pice, seed, 2)
    assert _rel(F2, Fo2) < tol
    assert _rel(F2, F) > 1e-5  # the shell part of the force is a few 1e-4 of the total


def test_rbl_with_rbe_md_graph_equals_eager(orc, gpu):
    """Random batch list + RBE k-space MD with list rebuilds (skin 0.1): graph replays and
    eager launches give the same bits, and the mean-force correction keeps the total
    momentum (the RBE k-space force also sums to zero)."""
    from copy_gpu_amd import systems

    n, steps = 3000, 24
    _log(f"case: random batch list + RBE MD, {n} ions, {steps} steps, skin 0.1")
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=SEED + 6)
    vx, vy, vz = systems.velocities(n, 1.0, seed=SEED + 6)
    Lb = np.array([L] * 3)
    outs = []
    for use_graph in (True, False):
        f = _make(Lb, n, 32, kspace="RBE", P=100, skin=0.1)
        f.Init(_dev(q))
        f.sample_kvectors(100, SEED, 0)
        X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
        F = [torch.zeros(n, device="cuda", dtype=torch.float32) for _ in range(3)]
        f.Execute(X, Y, Z, *F)
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=0.002, nsteps=steps, seed=SEED, step0=1,
                 use_graph=use_graph)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (X, Y, Z, VX, VY, VZ, *F)])
        builds = f.stats()["builds"]
        f.close()
    _log(f"  list builds {builds}")
    assert builds >= 3
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    p = np.array([o.sum() for o in outs[0][3:6]])
    p0 = np.array([vx.sum(), vy.sum(), vz.sum()])
    assert np.abs(p - p0).max() < 1e-4 * np.sqrt(n)
    assert np.all(np.isfinite(outs[0][0]))


@pytest.mark.parametrize("L,dims", [(10.0, 1), (12.0, 2)])
def test_small_box_images(orc, gpu, L, dims):
    """Boxes one and two list radii wide (list radius 5.3): the 27-cell stencil over
    one or two cells per axis lists every image within the radius (an atom once per
    image), and FP64 forces equal the oracle's (the cutoff 5.0 is at most L/2, so the
    minimum image is the only image within it)."""
    from copy_gpu_amd import systems

    n = int(0.5 * L ** 3)
    _log(f"case: small box L = {L} ({dims} cells per axis), {n} ions")
    x, y, z, q, Lb_ = systems.random_ions(n, n / L ** 3, 0.8, seed=SEED + int(L))
    Lb = np.array([L] * 3)
    f = _make(Lb, n, 64, rbl=False)
    f.Init(_dev(q))
    F = _compute(f, x, y, z)
    assert f.stats()["cells"] == dims ** 3
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    oo, io = orc.neighbors_images(xw, yw, zw, Lb, 5.3)
    og, ig = f.neighbors()
    assert og[-1] == oo[-1]
    for i in range(n):
        assert np.array_equal(np.sort(ig[og[i]:og[i + 1]]), np.sort(io[oo[i]:oo[i + 1]])), i
    Fo, evo = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5)
    assert _rel(F, Fo) < 1e-12
    ev = f.energy_virial()
    assert abs(ev["e_real"] - evo[1]) < 1e-11 * abs(evo[1])


def test_reference_example_input_runs(orc, gpu):
    """The reference's own example input (res/json_files/LJ_salt/rbmd.json, copied as a
    fixture): the in-built 10 x 10 x 10 lattice in a 10^3 box, cut_off 5 (one cell per
    axis), random batch list r_core 2.5 with neighbor_sample_num 50, RBE with P = 100 and
    alpha 0.1, Berendsen NVT.  Forces of one step against the oracle's random batch
    list + RBE forces on the same k-list, then 50 thermostatted steps and thermo rows."""
    from copy_gpu_amd import LJCutCoulKspaceForce, formats, systems, thermo

    cfg, run, _ = formats.load_config(REF_JSON)
    assert cfg.neighbor_type == "RBL" and cfg.r_core == 2.5 and cfg.rbl_pice == 10
    assert cfg.kspace == "RBE" and cfg.coulomb_sample_num == 100 and cfg.alpha == 0.1
    x, y, z, q = formats.inbuild_system(run, cfg.box_lo, cfg.box_len)
    n = x.size
    cfg.n_max = n
    cfg.rbl_seed = 5
    f = LJCutCoulKspaceForce(cfg)
    f.Init(_dev(q))
    Lb = np.array(cfg.box_len)
    P = cfg.coulomb_sample_num
    m, S = systems.sample_kvectors_host(cfg.alpha, Lb, P, seed=SEED, step=0)
    f.set_kvectors(_dev(m.reshape(-1), torch.int32))
    F = _compute(f, x, y, z)
    assert f.stats()["cells"] == 1
    epoch = f.stats()["builds"]
    lo = np.array(cfg.box_lo)
    Fr, _ = orc.rbl_forces(x, y, z, q, lo, Lb, 5.0, 5.0, cfg.alpha, 2.5, cfg.rbl_pice, 5,
                           epoch)
    Fk, _ = orc.kspace_forces(x, y, z, q, Lb, cfg.alpha, m, np.full(P, S / P))
    err = _rel(F, Fr + Fk)
    _log(f"  example input: {n} atoms, FP32 force rel L2 error {err:.3e}")
    assert err < 1e-5
    # the example's run: Berendsen NVT, temperature [1.0, 1.0, 0.5], timestep 2e-3
    vx, vy, vz = systems.velocities(n, 1.0, seed=SEED)
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    Fd = [torch.zeros(n, device="cuda", dtype=torch.float32) for _ in range(3)]
    f.sample_kvectors(P, SEED, 0)
    f.Execute(X, Y, Z, *Fd)
    f.set_thermostat(run["temp_ctrl_type"], run["temperature"])
    rows = []
    for k in range(5):
        f.md_run(X, Y, Z, VX, VY, VZ, *Fd, dt=run["dt"], nsteps=10, seed=SEED,
                 step0=1 + 10 * k)
        rows.append(thermo.thermo_row(f, X, Y, Z, 10 * (k + 1), run["dt"],
                                      rows[-1] if rows else None))
    torch.cuda.synchronize()
    T = [r["Temperature"] for r in rows]
    _log(f"  temperatures {['%.3f' % t for t in T]}")
    assert all(np.isfinite(r["TotalEnergy"]) for r in rows)
    assert all(0.2 < t < 5.0 for t in T)
