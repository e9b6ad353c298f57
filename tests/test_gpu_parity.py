"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on identical inputs.

Bars: neighbour lists equal as sets, bit for bit; forces within 1e-5 relative
(L2 over all atoms) in FP32 mode and 1e-12 in FP64 mode; structure factors and
energies/virial to the same relative bars.
"""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SEED = 20261202


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _make(Lb, n, precision=32, kspace="NONE", P=100, rc_lj=2.5, rc_coul=5.0, alpha=0.5,
          skin=0.3, lo=(0.0, 0.0, 0.0), eps=(1.0,), sigma=(1.0,), list_capacity=0,
          ewald_tol=1e-16, tile_stage=0, half_shell=False):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    cfg = ForceConfig(box_len=tuple(float(v) for v in Lb), box_lo=lo, cut_off=rc_lj,
                      coulomb_cut=rc_coul, skin=skin, alpha=alpha, kspace=kspace,
                      coulomb_sample_num=P, precision=precision, eps=eps, sigma=sigma,
                      n_max=n, list_capacity=list_capacity, ewald_tol=ewald_tol,
                      tile_stage=tile_stage, half_shell=half_shell)
    return LJCutCoulKspaceForce(cfg)


def _compute(f, x, y, z, ev=False):
    n = x.size
    X, Y, Z = _dev(x), _dev(y), _dev(z)
    F = [torch.zeros(n, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F, energy_virial=ev)
    torch.cuda.synchronize()
    return np.stack([t.double().cpu().numpy() for t in F], 1)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _sorted_rows(offs, idx):
    return [np.sort(idx[offs[i]:offs[i + 1]]) for i in range(len(offs) - 1)]


def _ions(n, seed=SEED, rho=0.3):
    from copy_gpu_amd import systems

    return systems.random_ions(n, rho, 0.8, seed=seed)


# ----------------------------------------------------------------------------- lists

def _check_sets(orc, f, x, y, z, lo, Lb, r_list):
    xw, yw, zw = orc.wrap(x, y, z, lo, Lb)
    oo, io = orc.neighbors_cells(xw, yw, zw, lo, Lb, r_list)
    og, ig = f.neighbors()
    assert og[-1] == oo[-1], (og[-1], oo[-1])
    ro, rg = _sorted_rows(oo, io), _sorted_rows(og, ig)
    bad = [i for i in range(len(ro)) if not np.array_equal(ro[i], rg[i])]
    assert not bad, f"{len(bad)} rows differ, first {bad[:5]}"


def test_neighbor_sets_random_ions(orc, gpu):
    _log("case: neighbor sets, 4000 random ions, density 0.3")
    x, y, z, q, L = _ions(4000)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size)
    f.Init(_dev(q))
    _compute(f, x, y, z)
    _check_sets(orc, f, x, y, z, np.zeros(3), Lb, 5.3)
    st = f.stats()
    assert st["builds"] == 1 and st["overflow_rows"] == 0
    # unwrapped input (atoms moved by whole periods) gives the same sets
    rng = np.random.default_rng(1)
    sh = rng.integers(-2, 3, size=(3, x.size)) * L
    f2 = _make(Lb, x.size)
    f2.Init(_dev(q))
    _compute(f2, x + sh[0], y + sh[1], z + sh[2])
    _check_sets(orc, f2, x + sh[0], y + sh[1], z + sh[2], np.zeros(3), Lb, 5.3)


def test_neighbor_sets_fcc_and_reference_fixture(orc, gpu):
    from copy_gpu_amd import systems

    _log("case: neighbor sets, FCC 4000 (density 0.8442)")
    x, y, z, L = systems.fcc(10, 0.8442)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, rc_coul=0.0)
    f.Init(_dev(np.zeros(x.size)))
    _compute(f, x, y, z)
    og, _ = f.neighbors()
    assert np.all(np.diff(og) == 78)
    _check_sets(orc, f, x, y, z, np.zeros(3), Lb, 2.8)

    _log("case: neighbor sets, reference LJ_salt 6000 ions, box 10x20x30")
    d = np.load(os.path.join(GOLDEN, "lj_salt_6000.npz"))
    Lb = d["hi"] - d["lo"]
    f = _make(Lb, 6000, rc_coul=2.5, lo=tuple(d["lo"]))
    f.Init(_dev(d["q"]))
    _compute(f, d["x"], d["y"], d["z"])
    _check_sets(orc, f, d["x"], d["y"], d["z"], d["lo"], Lb, 2.8)


def test_neighbor_sets_ragged_box_offset_origin(orc, gpu):
    _log("case: neighbor sets, ragged box with offset origin")
    rng = np.random.default_rng(7)
    lo = np.array([-3.0, 2.5, 10.0])
    Lb = np.array([17.0, 23.5, 19.25])
    n = 3000
    x, y, z = (lo[d] + rng.uniform(0, Lb[d], n) for d in range(3))
    f = _make(Lb, n, rc_coul=4.0, lo=tuple(lo))
    f.Init(_dev(np.zeros(n)))
    _compute(f, x, y, z)
    _check_sets(orc, f, x, y, z, lo, Lb, 4.3)


def _list_boundary_pairs(rl2):
    """Atom pairs whose FP64 squared distance, in the reference's operation order, is
    each double from 3 below to 3 above the largest d2 that (rl2 - d2) > 1e-4 accepts.
    A is at (18, y0, z0) with y0, z0 on a grid of 12 (whole numbers below 64), B at
    dx ~ sqrt(rl2) along x plus dy, dz of a few 2^-24: xB - 18 is exact (Sterbenz), dx*dx
    is a multiple of 2^-48 in [16, 32), y0 + dy and z0 + dz are exact and so are their
    differences, and dy^2, dz^2 are small multiples of 2^-48, so both FMAs of
    d2 = fma(dz, dz, fma(dy, dy, dx dx)) are exact sums and every d2 near the boundary is
    reachable."""
    t = rl2 - 1e-4
    while not (rl2 - t) > 1e-4:
        t = np.nextafter(t, -np.inf)
    while (rl2 - np.nextafter(t, np.inf)) > 1e-4:
        t = np.nextafter(t, np.inf)
    want = [t]
    for _ in range(3):
        want.insert(0, np.nextafter(want[0], -np.inf))
        want.append(np.nextafter(want[-1], np.inf))
    found = {}
    xb = 18.0 + np.sqrt(want[0]) - 1e-13
    for _ in range(80):
        dx = xb - 18.0
        base = dx * dx
        for a in range(6):
            for b in range(6):
                d2 = base + (a * a + b * b) * 2.0 ** -48
                if d2 in want and d2 not in found:
                    found[d2] = (xb, a * 2.0 ** -24, b * 2.0 ** -24)
        xb = np.nextafter(xb, np.inf)
    assert sorted(found) == sorted(want), "boundary distances not all reachable"
    return t, [found[d] for d in want]


@pytest.mark.parametrize("half_shell", [False, True])
def test_list_criterion_at_its_exact_boundary(orc, gpu, half_shell):
    """The reference's list criterion (r_list^2 - d2) > 1e-4 at its exact edge: pairs whose
    squared distance is each double from t - 3 ulp to t + 3 ulp (t the largest d2 it
    accepts), in the full-row and the half-shell builder: the rows equal the oracle's,
    and exactly the four pairs at or below t are kept."""
    _log(f"case: list criterion at its boundary, half_shell={half_shell}")
    rl2 = 5.3 * 5.3
    t, pairs = _list_boundary_pairs(rl2)
    L = 72.0
    pos = []
    for k, (xb, dy, dz) in enumerate(pairs):
        gy, gz = 12.0 * (k % 4), 12.0 * (k // 4)
        pos.append((18.0, 18.0 + gy, 18.0 + gz))
        pos.append((xb, 18.0 + gy + dy, 18.0 + gz + dz))
    x, y, z = (np.array([p[d] for p in pos]) for d in range(3))
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, rc_coul=5.0, half_shell=half_shell)
    f.Init(_dev(np.zeros(x.size)))
    _compute(f, x, y, z)
    _check_sets(orc, f, x, y, z, np.zeros(3), Lb, 5.3)
    og, ig = f.neighbors()
    kept = [bool(og[2 * k + 1] > og[2 * k]) for k in range(len(pairs))]
    assert kept == [True] * 4 + [False] * 3, kept


# ----------------------------------------------------------------------------- pair pass

@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_pair_forces_and_energy(orc, gpu, precision, tol):
    _log(f"case: pair forces FP{precision}, 8000 ions")
    x, y, z, q, L = _ions(8000)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, precision=precision)
    f.Init(_dev(q))
    F = _compute(f, x, y, z, ev=True)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, evo = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5)
    err = _rel(F, Fo)
    _log(f"  rel L2 force error {err:.3e}")
    assert err < tol
    floor = 1e-3 * np.sqrt(np.mean(np.sum(Fo ** 2, 1)))
    per_atom = np.linalg.norm(F - Fo, axis=1) / np.maximum(np.linalg.norm(Fo, axis=1), floor)
    assert per_atom.max() < 50 * tol, per_atom.max()
    ev = f.energy_virial()
    assert abs(ev["e_lj"] - evo[0]) < 10 * tol * abs(evo[0])
    assert abs(ev["e_real"] - evo[1]) < 10 * tol * abs(evo[1])
    W = ev["virial"]
    assert np.linalg.norm(W - evo[2:8]) < 10 * tol * np.linalg.norm(evo[2:8])


@pytest.mark.parametrize("types", [False, True])
def test_half_shell_rows_equal_full_rows(orc, gpu, types):
    """FP32 handles with half_shell evaluate each pair once over half-shell tiles (home
    cell + 13 forward cells) and add the force to both atoms; with full rows every pair is
    evaluated from both sides.  Same neighbour sets, forces within FP32 rounding,
    energies and virial to 1e-6, both against the oracle at its FP32 bars; and over
    tiles read from global memory (LDS stage 64) the same bits as staged tiles."""
    _log(f"case: half-shell vs full rows, FP32, 12000 ions, two types {types}")
    x, y, z, q, L = _ions(12000, seed=31)
    Lb = np.array([L] * 3)
    kw = dict(eps=(1.0, 0.6), sigma=(1.0, 1.3)) if types else {}
    tt = (np.arange(x.size) % 3 == 0).astype(np.int32) if types else None
    out = {}
    for full, stage in ((False, 0), (True, 0), (False, 64)):
        f = _make(Lb, x.size, **kw, half_shell=not full, tile_stage=stage)
        f.Init(_dev(q), None if tt is None else _dev(tt, torch.int32))
        F = _compute(f, x, y, z, ev=True)
        out[(full, stage)] = (F, f.energy_virial(), f.neighbors(), f.stats())
        f.close()
    Fh, evh, (oh, ih), sh = out[(False, 0)]
    Ff, evf, (of, if_), sf = out[(True, 0)]
    assert sh["entries"] == sf["entries"]
    assert np.array_equal(oh, of)
    for i in range(0, x.size, 7):
        assert np.array_equal(np.sort(ih[oh[i]:oh[i + 1]]), np.sort(if_[of[i]:of[i + 1]]))
    assert _rel(Fh, Ff) < 2e-6, _rel(Fh, Ff)
    for k in ("e_lj", "e_real"):
        assert abs(evh[k] - evf[k]) < 1e-6 * abs(evf[k])
    assert np.linalg.norm(evh["virial"] - evf["virial"]) < 1e-6 * np.linalg.norm(evf["virial"])
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, evo = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5, types=tt,
                              eps=kw.get("eps", (1.0,)), sigma=kw.get("sigma", (1.0,)))
    assert _rel(Fh, Fo) < 1e-5
    # the global-memory tiles: the same bits (order-independent LDS sums)
    Fg, evg, _, sg = out[(False, 64)]
    assert sg["big_tiles"] > 0
    assert np.array_equal(Fg, Fh)
    assert evg["e_real"] == evh["e_real"]


def test_time_pair_leaves_a_whole_evaluation(gpu):
    """rbmd_force_time_pair: a positive mean per pair-pass launch, and the force arrays
    end holding the same bits as a plain compute at the same positions (pair plus
    k-space forces)."""
    _log("case: time_pair, 8000 ions, RBE P=100")
    x, y, z, q, L = _ions(8000)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, kspace="RBE")
    f.Init(_dev(q))
    f.sample_kvectors(100, SEED, 0)
    F0 = _compute(f, x, y, z)
    X, Y, Z = _dev(x), _dev(y), _dev(z)
    F = [torch.zeros(x.size, device="cuda", dtype=f.fdtype) for _ in range(3)]
    ms = f.time_pair(X, Y, Z, *F, reps=5)
    torch.cuda.synchronize()
    F1 = np.stack([t.double().cpu().numpy() for t in F], 1)
    assert 0.0 < ms < 100.0, ms
    assert np.array_equal(F1, F0)


def test_pair_forces_two_types(orc, gpu):
    _log("case: pair forces FP64, two LJ types")
    x, y, z, q, L = _ions(3000, seed=3)
    Lb = np.array([L] * 3)
    types = (np.arange(x.size) % 3 == 0).astype(np.int32)
    eps, sig = (1.0, 0.6), (1.0, 1.3)
    f = _make(Lb, x.size, precision=64, eps=eps, sigma=sig)
    f.Init(_dev(q), _dev(types, torch.int32))
    F = _compute(f, x, y, z)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, _ = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5, types=types,
                            eps=eps, sigma=sig)
    assert _rel(F, Fo) < 1e-12


def test_overflowed_rows_use_stencil_fallback(orc, gpu):
    _log("case: list capacity 64 (rows overflow), FP64")
    x, y, z, q, L = _ions(3000, seed=9)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, precision=64, list_capacity=64)
    f.Init(_dev(q))
    F = _compute(f, x, y, z)
    assert f.stats()["overflow_rows"] > 0
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, _ = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5)
    assert _rel(F, Fo) < 1e-12
    # the next call grows the rows and rebuilds
    F2 = _compute(f, x, y, z)
    st = f.stats()
    assert st["overflow_rows"] == 0 and st["cap"] > 64
    assert _rel(F2, Fo) < 1e-12


@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_rows_near_capacity_mix_fitted_and_overflowed(orc, gpu, precision, tol):
    """Capacity 192 against rows of 187 +- 14 entries: in most waves some rows fit
    (read up to their sentinel) and others overflowed (stencil scan), so the pair
    rounds take their bounds-tested path.  Forces, energies and virial must not
    change; the sentinel's dummy entry adds exactly zero."""
    _log(f"case: list capacity 192 (some rows overflow), precision {precision}")
    x, y, z, q, L = _ions(6000, seed=21)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, precision=precision, list_capacity=192)
    f.Init(_dev(q))
    F = _compute(f, x, y, z, ev=True)
    st = f.stats()
    assert 0 < st["overflow_rows"] < x.size, st
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, evo = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5)
    assert _rel(F, Fo) < tol
    ev = f.energy_virial()
    assert abs(ev["e_lj"] - evo[0]) < 10 * tol * abs(evo[0])
    assert abs(ev["e_real"] - evo[1]) < 10 * tol * abs(evo[1])


@pytest.mark.parametrize("stage", [64, 1216])
def test_tiles_beyond_lds_stage_read_global_memory(orc, gpu, stage):
    """Tiles with more candidates than the LDS stage take the global-memory
    paths of the list build and the pair pass: stage 64 sends every tile there,
    stage 1216 (about the mean tile of this system) splits the tiles between the
    two pair kernels.  Sets and forces must not change."""
    _log(f"case: LDS stage {stage} atoms, FP64, 6000 ions")
    x, y, z, q, L = _ions(6000, seed=13)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size, precision=64, tile_stage=stage)
    f.Init(_dev(q))
    F = _compute(f, x, y, z, ev=True)
    _check_sets(orc, f, x, y, z, np.zeros(3), Lb, 5.3)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fo, evo = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, 0.5)
    assert _rel(F, Fo) < 1e-12
    ev = f.energy_virial()
    assert abs(ev["e_real"] - evo[1]) < 1e-11 * abs(evo[1])


# ----------------------------------------------------------------------------- k-space

@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_rbe_forces_given_klist(orc, gpu, precision, tol):
    from copy_gpu_amd import systems

    _log(f"case: RBE FP{precision}, 32768 ions, P=100 (c2 size)")
    x, y, z, q, L = _ions(32768)
    Lb = np.array([L] * 3)
    alpha, P = 0.5, 100
    m, S = systems.sample_kvectors_host(alpha, Lb, P, seed=SEED, step=0)
    f = _make(Lb, x.size, precision=precision, kspace="RBE", P=P)
    f.Init(_dev(q))
    f.set_kvectors(_dev(m.reshape(-1), torch.int32))
    F = _compute(f, x, y, z, ev=True)
    mg, wg, re, im = f.kspace()
    assert np.array_equal(mg, m)
    assert np.allclose(wg, S / P, rtol=0, atol=0)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    reo, imo = orc.rho(xw, yw, zw, q, Lb, m)
    scale = np.max(np.hypot(reo, imo))
    rerr = max(np.abs(re - reo).max(), np.abs(im - imo).max()) / scale
    _log(f"  rho max error / max|rho| = {rerr:.3e}")
    assert rerr < tol
    Fp, evp = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, alpha)
    Fk, evk = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, m, np.full(P, S / P), reo, imo)
    err = _rel(F, Fp + Fk)
    kerr = _rel(F - Fp, Fk)
    _log(f"  rel L2 total force error {err:.3e}, k-space part {kerr:.3e}")
    assert err < tol
    assert kerr < 10 * tol
    ev = f.energy_virial()
    assert abs(ev["e_kspace"] - evk[0]) < 10 * tol * abs(evk[0])
    assert abs(ev["e_self"] - orc.self_energy(q, alpha)) < 1e-12 * abs(ev["e_self"])
    W = ev["virial"]
    Wo = evp[2:8] + evk[1:7]
    assert np.linalg.norm(W - Wo) < 10 * tol * np.linalg.norm(Wo)


@pytest.mark.parametrize("n,slices", [(20000, 8), (70000, 4), (140000, 2), (270000, 1)])
def test_kspace_force_work_splits(orc, gpu, n, slices):
    """The k-space force kernel splits a workgroup's 8 waves into atom groups of 64 x k
    slices by the launch size (8 slices below 65,409 atoms at P <= 128, fewer above, so
    it still has 512 workgroups): every split gives the oracle's forces on the same k-list
    (FP32, the k-space part's bar 1e-4 relative L2; a wrong slice sum is off by O(1))."""
    from copy_gpu_amd import force as F
    from copy_gpu_amd import systems

    P, alpha = 64, 0.5
    _log(f"case: k-space force, {n} ions, P={P}, expected {slices} slices per atom group")
    x, y, z, q, L = _ions(n, seed=SEED + n)
    Lb = np.array([L] * 3)
    m, S = systems.sample_kvectors_host(alpha, Lb, P, seed=SEED, step=3)
    m = np.asarray(m, np.int32).reshape(-1, 3)
    fx, fy, fz = (torch.zeros(n, device="cuda", dtype=torch.float32) for _ in range(3))
    F.kspace_compute(Lb, alpha, _dev(x), _dev(y), _dev(z), _dev(q),
                     _dev(m.reshape(-1), torch.int32), None, fx, fy, fz, precision=32)
    torch.cuda.synchronize()
    Fg = torch.stack([fx, fy, fz], 1).double().cpu().numpy()
    Fo, _ = orc.kspace_forces(x, y, z, q, Lb, alpha, m, np.full(P, S / P))
    err = _rel(Fg, Fo)
    _log(f"  rel L2 error {err:.3e}")
    assert err < 1e-4


def test_autotuned_shapes_keep_parity(orc, gpu):
    """The autotuner only changes launch shapes (workgroups per tile, LDS stage, rho(k)
    workgroup size): forces with the tuned shapes still meet the FP32 bar."""
    from copy_gpu_amd import systems

    _log("case: autotune, 12000 ions, RBE FP32 P=100, then parity against the oracle")
    x, y, z, q, L = _ions(12000, seed=77)
    Lb = np.array([L] * 3)
    alpha, P = 0.5, 100
    m, S = systems.sample_kvectors_host(alpha, Lb, P, seed=SEED, step=4)
    f = _make(Lb, x.size, kspace="RBE", P=P)
    f.Init(_dev(q))
    f.set_kvectors(_dev(m.reshape(-1), torch.int32))
    X, Y, Z = _dev(x), _dev(y), _dev(z)
    F = [torch.zeros(x.size, device="cuda", dtype=torch.float32) for _ in range(3)]
    shapes = f.autotune(X, Y, Z, *F, reps=2)
    _log(f"  tuned shapes {shapes}")
    assert shapes["tuned_cost_us"] > 0
    assert 1 <= shapes["split"] <= 4 and shapes["rho_atoms_per_workgroup"] % 64 == 0
    Ft = torch.stack(F, 1).double().cpu().numpy()
    Fc = _compute(f, x, y, z)
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    reo, imo = orc.rho(xw, yw, zw, q, Lb, m)
    Fp, _ = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, alpha)
    Fk, _ = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, m, np.full(P, S / P), reo, imo)
    assert _rel(Fc, Fp + Fk) < 1e-5
    assert _rel(Ft, Fp + Fk) < 1e-5


def test_device_sampler_matches_host(gpu):
    from copy_gpu_amd import systems

    _log("case: device k-sampler vs host draws")
    L = 69.336
    Lb = np.array([L] * 3)
    f = _make(Lb, 10, kspace="RBE", P=400, alpha=0.5)
    for step in (0, 1, 977):
        f.sample_kvectors(400, SEED, step)
        m, w, _, _ = f.kspace()
        mh, S = systems.sample_kvectors_host(0.5, Lb, 400, SEED, step)
        assert np.array_equal(m, mh)
        assert np.all(w == S / 400)


def _halfspace_lattice(Lb, alpha, tol):
    kmax2 = -4 * alpha * np.log(tol)
    M = [int(np.floor(np.sqrt(kmax2) * Lb[d] / (2 * np.pi))) + 1 for d in range(3)]
    a, b, g = np.meshgrid(np.arange(0, M[0] + 1), np.arange(-M[1], M[1] + 1),
                          np.arange(-M[2], M[2] + 1), indexing="ij")
    a, b, g = a.ravel(), b.ravel(), g.ravel()
    keep = ~((a == 0) & ((b < 0) | ((b == 0) & (g <= 0))))
    a, b, g = a[keep], b[keep], g[keep]
    k2 = (2 * np.pi * a / Lb[0]) ** 2 + (2 * np.pi * b / Lb[1]) ** 2 + (2 * np.pi * g / Lb[2]) ** 2
    keep = k2 <= kmax2
    m = np.stack([a[keep], b[keep], g[keep]], 1).astype(np.int32)
    return m, 2 * np.exp(-k2[keep] / (4 * alpha))


def test_ewald_mode_fp64(orc, gpu):
    _log("case: full Ewald FP64, 2000 ions")
    x, y, z, q, L = _ions(2000, seed=21)
    Lb = np.array([L] * 3)
    alpha, tol_k = 0.5, 1e-12
    f = _make(Lb, x.size, precision=64, kspace="EWALD", alpha=alpha, ewald_tol=tol_k)
    f.Init(_dev(q))
    F = _compute(f, x, y, z, ev=True)
    m, w = _halfspace_lattice(Lb, alpha, tol_k)
    assert f.stats()["ewald_k"] == m.shape[0]
    xw, yw, zw = orc.wrap(x, y, z, np.zeros(3), Lb)
    Fp, evp = orc.pair_forces(xw, yw, zw, q, np.zeros(3), Lb, 2.5, 5.0, alpha)
    Fk, evk = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, m, w)
    assert _rel(F, Fp + Fk) < 1e-12
    # physics: the half-space sum equals the full-lattice Ewald sum
    kc = [int(np.ceil(np.sqrt(-4 * alpha * np.log(tol_k)) * L / (2 * np.pi))) + 1] * 3
    mf, wf = orc.ewald_kvectors(Lb, alpha, kc)
    Ff, evf = orc.kspace_forces(xw, yw, zw, q, Lb, alpha, mf, wf)
    assert _rel(F - Fp, Ff) < 1e-9
    ev = f.energy_virial()
    assert abs(ev["e_kspace"] - evf[0]) < 1e-9 * abs(evf[0])


# ----------------------------------------------------------------------------- MD loop

def _state(n, seed=SEED):
    from copy_gpu_amd import systems

    x, y, z, q, L = _ions(n, seed=seed)
    vx, vy, vz = systems.velocities(n, 1.0, seed=seed)
    return x, y, z, q, L, vx, vy, vz


def test_md_graph_equals_eager(gpu):
    _log("case: 40 MD steps (a 32-step graph and an 8-step one), graph replay vs eager")
    x, y, z, q, L, vx, vy, vz = _state(4000)
    Lb = np.array([L] * 3)
    outs = []
    for use_graph in (False, True):
        f = _make(Lb, x.size, kspace="RBE", P=100)
        f.Init(_dev(q))
        f.sample_kvectors(100, SEED, 0)
        X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
        F = [torch.zeros(x.size, device="cuda", dtype=torch.float32) for _ in range(3)]
        f.Execute(X, Y, Z, *F)
        img = torch.zeros(3 * x.size, device="cuda", dtype=torch.int32)
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=0.002, nsteps=40, seed=SEED, step0=1, img=img,
                 use_graph=use_graph)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (X, Y, Z, VX, VY, VZ, *F, img)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_md_matches_oracle_integrator(orc, gpu):
    """FP64 velocity Verlet with RBE (device-sampled k per step) against the oracle
    loop kick / drift / forces / kick on the same k-lists."""
    from copy_gpu_amd import systems

    _log("case: 10 MD steps FP64 RBE vs oracle loop, 2000 ions")
    n = 2000
    x, y, z, q, L, vx, vy, vz = _state(n, seed=31)
    Lb = np.array([L] * 3)
    lo = np.zeros(3)
    alpha, P, dt, steps = 0.5, 50, 0.002, 10
    f = _make(Lb, n, precision=64, kspace="RBE", P=P)
    f.Init(_dev(q))
    f.sample_kvectors(P, SEED, 0)
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1, use_graph=True)
    torch.cuda.synchronize()

    def forces(px, py, pz, step):
        m, S = systems.sample_kvectors_host(alpha, Lb, P, SEED, step)
        Fp, _ = orc.pair_forces(px, py, pz, q, lo, Lb, 2.5, 5.0, alpha)
        Fk, _ = orc.kspace_forces(px, py, pz, q, Lb, alpha, m, np.full(P, S / P))
        return Fp + Fk

    p = list(orc.wrap(x, y, z, lo, Lb))
    v = [vx, vy, vz]
    Fo = forces(*p, 0)
    for s in range(1, steps + 1):
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
        p, _ = orc.drift(p, v, dt, lo, Lb)
        Fo = forces(*p, s)
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
    Pg = np.stack([X.cpu().numpy(), Y.cpu().numpy(), Z.cpu().numpy()], 1)
    Po = np.stack(p, 1)
    dpos = Pg - Po
    dpos -= Lb * np.rint(dpos / Lb)
    assert np.abs(dpos).max() < 1e-9, np.abs(dpos).max()
    Vg = np.stack([VX.cpu().numpy(), VY.cpu().numpy(), VZ.cpu().numpy()], 1)
    assert np.abs(Vg - np.stack(v, 1)).max() < 1e-8


@pytest.mark.parametrize("n,chunked", [(4000, False), (40000, True)])
def test_md_many_cells_chunked_cell_sort(orc, gpu, n, chunked):
    """Cells of about one atom (cutoffs 1.0, skin 0.1): 9,261 cells at 4,000 ions and
    97,336 at 40,000; above 16,384 cells the MD cell sort takes each cell's start from
    256-cell chunk counts kept beside the cell counts.  FP64 LJ + real-space Coulomb MD
    with several rebuilds against the oracle loop (the hot start moves atoms up to
    0.023 per step, under half the skin)."""
    dt, steps = 0.002, 12
    x, y, z, q, L, vx, vy, vz = _state(n, seed=47)
    Lb = np.array([L] * 3)
    lo = np.zeros(3)
    _log(f"case: {n} ions, cutoffs 1.0, skin 0.1, {steps} FP64 MD steps with rebuilds")
    f = _make(Lb, n, precision=64, kspace="NONE", rc_lj=1.0, rc_coul=1.0, skin=0.1)
    f.Init(_dev(q))
    assert (f.stats()["cells"] > 16384) == chunked
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    b0 = f.stats()["builds"]
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=dt, nsteps=steps, seed=SEED, step0=1, use_graph=True)
    torch.cuda.synchronize()
    assert f.stats()["builds"] - b0 >= 2

    def forces(px, py, pz):
        Fp, _ = orc.pair_forces(px, py, pz, q, lo, Lb, 1.0, 1.0, 0.5)
        return Fp

    p = list(orc.wrap(x, y, z, lo, Lb))
    v = [vx, vy, vz]
    Fo = forces(*p)
    for s in range(1, steps + 1):
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
        p, _ = orc.drift(p, v, dt, lo, Lb)
        Fo = forces(*p)
        v = orc.kick(v, [Fo[:, 0], Fo[:, 1], Fo[:, 2]], dt)
    Pg = np.stack([X.cpu().numpy(), Y.cpu().numpy(), Z.cpu().numpy()], 1)
    dpos = Pg - np.stack(p, 1)
    dpos -= Lb * np.rint(dpos / Lb)
    bad = np.nonzero(np.abs(dpos).max(1) > 1e-9)[0]
    _log(f"  builds {f.stats()['builds'] - b0}, atoms off by > 1e-9: {bad.size} {bad[:8]}")
    assert np.abs(dpos).max() < 1e-9, np.abs(dpos).max()
    Vg = np.stack([VX.cpu().numpy(), VY.cpu().numpy(), VZ.cpu().numpy()], 1)
    assert np.abs(Vg - np.stack(v, 1)).max() < 1e-8


def test_md_step_longer_than_half_skin_is_an_error(gpu):
    """A rebuild takes its reference positions from the binning one step earlier, so
    an MD step that moves an atom more than half the skin can leave a pair out of the
    list: the run reports it.  The hot start of 4,000 ions moves an atom 0.0195 in one
    step (oracle loop), past half of a 0.02 skin."""
    from copy_gpu_amd import RBMDError

    n = 4000
    x, y, z, q, L, vx, vy, vz = _state(n, seed=47)
    Lb = np.array([L] * 3)
    _log("case: MD with a skin shorter than twice the largest step (error expected)")
    f = _make(Lb, n, precision=64, kspace="NONE", rc_lj=1.0, rc_coul=1.0, skin=0.02)
    f.Init(_dev(q))
    X, Y, Z, VX, VY, VZ = (_dev(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    with pytest.raises(RBMDError, match="half the Verlet skin"):
        f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=0.002, nsteps=12, seed=SEED, step0=1)
    torch.cuda.synchronize()


def test_rebuild_trigger(gpu):
    _log("case: list rebuild only after half-skin displacement")
    x, y, z, q, L = _ions(2000, seed=41)
    Lb = np.array([L] * 3)
    f = _make(Lb, x.size)
    f.Init(_dev(q))
    _compute(f, x, y, z)
    assert f.stats()["builds"] == 1
    _compute(f, x + 0.1, y, z)  # 0.1 < skin/2 = 0.15
    assert f.stats()["builds"] == 1
    x2 = x.copy()
    x2[5] += 0.16
    _compute(f, x2, y, z)
    assert f.stats()["builds"] == 2


def test_errors_and_empty(gpu):
    from copy_gpu_amd import RBMDError

    _log("case: error paths")
    # a box narrower than one list radius (5.3) on some axis: the stencil would miss
    # images (boxes of one or two list radii are supported, test_gpu_rbl)
    with pytest.raises(RBMDError, match="one list radius"):
        _make(np.array([10.0, 5.0, 10.0]), 10)
    with pytest.raises(RBMDError, match="r_core"):
        from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

        LJCutCoulKspaceForce(ForceConfig(box_len=(20.0,) * 3, n_max=10, neighbor_type="RBL",
                                         r_core=6.0, rbl_pice=3))
    f = _make(np.array([20.0] * 3), 100, kspace="RBE", P=10)
    with pytest.raises(RBMDError, match="set_atoms"):
        _compute(f, np.zeros(0), np.zeros(0), np.zeros(0))
    with pytest.raises(RBMDError, match="P out of"):
        f.sample_kvectors(11, 1, 0)
    f.Init(_dev(np.zeros(1)))
    with pytest.raises(RBMDError, match="no k-vectors"):
        _compute(f, np.ones(1), np.ones(1), np.ones(1))
    g = _make(np.array([20.0] * 3), 100)
    g.Init(_dev(np.zeros(0)))
    _compute(g, np.zeros(0), np.zeros(0), np.zeros(0))  # empty system is a no-op
    with pytest.raises(RBMDError, match="n exceeds"):
        g.Init(_dev(np.zeros(101)))
    # rows so long that the list build's per-wave hit buffers cannot fit the LDS
    with pytest.raises(RBMDError, match="do not fit the LDS"):
        _make(np.array([20.0] * 3), 100, list_capacity=16384)


def test_million_ions_invariants(gpu):
    """c3 size (1,048,576 ions, alpha 0.3, rc 6.0, P 500): total force sums to zero
    (pair part antisymmetric, RBE part sum_l k_l (Re Im - Im Re) = 0) and forces
    are invariant under a common translation of all atoms."""
    _log("case: 1,048,576 ions, RBE P=500, invariants")
    n = 1 << 20
    x, y, z, q, L = _ions(n)
    Lb = np.array([L] * 3)
    f = _make(Lb, n, kspace="RBE", P=500, alpha=0.3, rc_coul=6.0)
    f.Init(_dev(q))
    f.sample_kvectors(500, SEED, 0)
    F = _compute(f, x, y, z)
    scale = np.linalg.norm(F) / np.sqrt(n)
    assert np.abs(F.sum(0)).max() < 1e-3 * scale * np.sqrt(n)
    F2 = _compute(f, x + 1.2345, y - 3.5, z + 0.25)
    assert _rel(F2, F) < 2e-5
    st = f.stats()
    assert st["overflow_rows"] == 0
    mean_row = st["entries"] / n
    assert abs(mean_row - 0.3 * 4 / 3 * np.pi * 6.3 ** 3) < 5
