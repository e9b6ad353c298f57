"""GPU edge cases through the C ABI: an empty system, and two ions whose force,
energies and virial have closed forms (LJ 12-6 plus the real-space Ewald term
erfc(sqrt(alpha) r) / r, the pair terms of ExecForceFunction.h:74-96 and 217-224),
inside both cutoffs, inside the erfc cutoff only, in the skin beyond both, and
across the periodic boundary."""
import math
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ALPHA, RC_LJ, RC_COUL, L = 0.5, 2.5, 5.0, 20.0


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _make(n_max, precision):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    cfg = ForceConfig(box_len=(L, L, L), cut_off=RC_LJ, coulomb_cut=RC_COUL, skin=0.3,
                      alpha=ALPHA, kspace="NONE", precision=precision, n_max=n_max)
    return LJCutCoulKspaceForce(cfg)


def _closed_form(r, qq):
    """Force factor fs (force on ion 0 = fs (r_1 - r_0)), E_LJ, E_real of one pair."""
    fs = e_lj = e_re = 0.0
    if r * r < RC_LJ * RC_LJ:
        sr6 = r ** -6
        fs += -24.0 * (2.0 * sr6 - 1.0) * sr6 / (r * r)
        e_lj = 4.0 * (sr6 * sr6 - sr6)
    if r * r < RC_COUL * RC_COUL:
        ec = math.erfc(math.sqrt(ALPHA) * r)
        g = ec / (r * r) + 2.0 * math.sqrt(ALPHA / math.pi) * math.exp(-ALPHA * r * r) / r
        fs -= qq * g / r
        e_re = qq * ec / r
    return fs, e_lj, e_re


@pytest.mark.parametrize("precision", [64, 32])
def test_empty_system(gpu, precision):
    _log(f"case: empty system, precision {precision}")
    f = _make(16, precision)
    f.Init(torch.zeros(0, device="cuda", dtype=torch.float64))
    e = torch.zeros(0, device="cuda", dtype=torch.float64)
    F = [torch.zeros(0, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(e, e, e, *F, energy_virial=True)
    v = [torch.zeros(0, device="cuda", dtype=torch.float64) for _ in range(3)]
    f.md_run(e, e, e, *v, *F, dt=0.002, nsteps=5, seed=1, step0=1)
    torch.cuda.synchronize()
    f.close()


# (ion 0, ion 1): separations inside both cutoffs, the erfc cutoff only, the skin
# beyond both (zero force), and 1.2 across the x boundary
PAIRS = [((5.0, 5.0, 5.0), (6.3, 5.0, 5.0)),
         ((5.0, 5.0, 5.0), (5.0, 8.1, 6.6)),
         ((5.0, 5.0, 5.0), (5.0, 5.0, 10.2)),
         ((0.4, 9.0, 9.0), (19.2, 9.0, 9.0))]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("k", range(len(PAIRS)))
def test_two_ions_closed_form(gpu, precision, k):
    p0, p1 = (np.array(p) for p in PAIRS[k])
    d = p1 - p0
    d -= L * np.round(d / L)  # minimum image
    r = float(np.linalg.norm(d))
    q = np.array([1.0, -1.0])
    fs, e_lj, e_re = _closed_form(r, q[0] * q[1])
    _log(f"case: two ions at r = {r:.4f}, precision {precision}")
    f = _make(2, precision)
    f.Init(_dev(q))
    F = [torch.zeros(2, device="cuda", dtype=f.fdtype) for _ in range(3)]
    X, Y, Z = (_dev([p0[a], p1[a]]) for a in range(3))
    f.Execute(X, Y, Z, *F, energy_virial=True)
    torch.cuda.synchronize()
    got = np.stack([t.double().cpu().numpy() for t in F], 1)
    want = np.stack([fs * d, -fs * d])
    ev = f.energy_virial()
    f.close()
    if fs == 0.0:
        assert np.all(got == 0.0), got
        assert ev["e_lj"] == 0.0 and ev["e_real"] == 0.0
        return
    # virial W_ab = -fs d_a d_b (each ion's half of the pair)
    W = -fs * np.array([d[0] * d[0], d[1] * d[1], d[2] * d[2], d[0] * d[1], d[0] * d[2],
                        d[1] * d[2]])
    if precision == 64:
        rel, erfc_abs = 1e-12, 0.0
    else:
        # FP32 pair maths: erfc by Abramowitz & Stegun 7.1.26, |error| <= 1.5e-7
        # absolute, which moves E_real by up to 1.5e-7 |qq| / r, each force component
        # by up to 1.5e-7 |qq| / r^2 and the virial by up to 1.5e-7 |qq| / r
        rel, erfc_abs = 1e-5, 1.5e-7 * abs(q[0] * q[1])
    assert np.abs(got - want).max() <= rel * np.abs(want).max() + erfc_abs / r ** 2, (got, want)
    assert abs(ev["e_lj"] - e_lj) <= rel * abs(e_lj)
    assert abs(ev["e_real"] - e_re) <= rel * abs(e_re) + erfc_abs / r
    assert np.abs(ev["virial"] - W).max() <= rel * np.abs(W).max() + erfc_abs / r, (
        ev["virial"], W)


def test_more_atoms_than_capacity_is_an_error(gpu):
    from copy_gpu_amd import RBMDError

    _log("case: set_atoms beyond n_max (error expected)")
    f = _make(4, 64)
    with pytest.raises(RBMDError, match="n_max"):
        f.Init(_dev(np.ones(5)))
    f.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_lone_ion_feels_no_force(gpu, precision):
    """A single charge in the periodic box: no pair partner, and its k-space force
    (the sum over k of the ion's own structure factor against its own phase) vanishes,
    so no self-force leaks through rho(k)."""
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    _log(f"case: one ion, RBE P = 200, precision {precision}")
    cfg = ForceConfig(box_len=(L, L, L), cut_off=RC_LJ, coulomb_cut=RC_COUL, skin=0.3,
                      alpha=ALPHA, kspace="RBE", coulomb_sample_num=200, precision=precision,
                      n_max=1)
    f = LJCutCoulKspaceForce(cfg)
    f.Init(_dev([2.0]))
    f.sample_kvectors(200, 7, 0)
    F = [torch.zeros(1, device="cuda", dtype=f.fdtype) for _ in range(3)]
    f.Execute(_dev([3.3]), _dev([11.7]), _dev([17.2]), *F)
    torch.cuda.synchronize()
    got = np.array([float(t.cpu()[0]) for t in F])
    f.close()
    # one k-vector's term has magnitude of order q^2 S/(P V) |k| <= 4 / 8000 * 2.4; FP32
    # phases leave a cancellation residue of a few ulp of that per term
    assert np.abs(got).max() <= (1e-12 if precision == 64 else 1e-6), got
