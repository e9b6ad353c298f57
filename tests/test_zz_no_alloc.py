"""rbmd_kspace_compute_noalloc (fixed per-device scratch, no allocation) against
rbmd_kspace_compute (scratch allocated per call): the same kernels and geometry, so
forces, rho(k) and energy / virial must agree bit for bit.  Two calls in a row, the
second with fewer atoms and k-vectors, so the second call runs over scratch that
the first one left filled.  Both entry points are also checked against the oracle's
k-space forces (relative L2: FP64 1e-12; FP32 1e-4, the bar of the k-space part of
the FP32 force tests).
"""
import numpy as np
import pytest
import torch

from copy_gpu_amd import force as F
from copy_gpu_amd import systems

pytestmark = pytest.mark.gpu


def _case(n, P, seed, learned):
    x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=seed)
    Lb = np.array([L] * 3)
    m, S = systems.sample_kvectors_host(0.5, Lb, P, seed=seed, step=0)
    m = np.asarray(m, np.int32).reshape(-1, 3)
    if learned:
        # importance weights of an arbitrary positive proposal, as a learned run has
        w = np.random.default_rng(seed).uniform(0.5, 1.5, P) * (S / P)
    else:
        w = None
    return x, y, z, q, Lb, m, w, S


def _run(gpu, case, precision, noalloc):
    x, y, z, q, Lb, m, w, _ = case
    dev = lambda a, t=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(gpu, t)  # noqa: E731
    dt = torch.float32 if precision == 32 else torch.float64
    fx, fy, fz = (torch.zeros(x.size, device=gpu, dtype=dt) for _ in range(3))
    rho, ev = F.kspace_compute(Lb, 0.5, dev(x), dev(y), dev(z), dev(q),
                               dev(m.reshape(-1), torch.int32),
                               None if w is None else dev(w), fx, fy, fz,
                               precision=precision, noalloc=noalloc)
    torch.cuda.synchronize()
    Fk = torch.stack([fx, fy, fz], 1).cpu().numpy()
    return Fk, rho.cpu().numpy(), ev.cpu().numpy()


@pytest.mark.parametrize("precision", [32, 64])
def test_noalloc_matches_allocating_entry_point(orc, gpu, precision):
    tol = 1e-4 if precision == 32 else 1e-12
    cases = [_case(20000, 300, 11, learned=True), _case(3000, 70, 12, learned=False)]
    for k, case in enumerate(cases):
        print(f"noalloc case {k}: n={case[0].size} P={case[5].shape[0]} precision={precision}",
              flush=True)
        Fa, ra, ea = _run(gpu, case, precision, noalloc=False)
        Fb, rb, eb = _run(gpu, case, precision, noalloc=True)
        assert np.array_equal(Fa, Fb)
        assert np.array_equal(ra, rb)
        assert np.array_equal(ea, eb)
        x, y, z, q, Lb, m, w, S = case
        wv = np.full(m.shape[0], S / m.shape[0]) if w is None else w
        Fo, evo = orc.kspace_forces(x, y, z, q, Lb, 0.5, m, wv)
        err = np.linalg.norm(Fb - Fo) / np.linalg.norm(Fo)
        assert err < tol, err
        assert abs(eb[0] - evo[0]) <= 10 * tol * abs(evo[0])


def test_noalloc_rejects_p_beyond_scratch(gpu):
    case = _case(100, 8193, 13, learned=False)
    with pytest.raises(F.RBMDError, match="8192"):
        _run(gpu, case, 64, noalloc=True)
