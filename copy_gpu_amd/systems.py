"""Deterministic synthetic inputs (host, numpy) from the library's Philox generators.

Configurations (reduced LJ units):
  fcc_lj      N = 4 nc^3 atoms on an FCC lattice at density rho (no charges)
  electrolyte N ions uniform in a cubic box at density rho, minimum separation 0.8,
              half +1 half -1 by a seeded shuffle
  replicas    sizes N_r = floor(n_min (n_max / n_min)^u_r), u_r uniform from Philox
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

DEFAULT_SEED = 20261202


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def philox4x32(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    L.lib().rbmd_philox4x32(_p(c), _p(k), _p(out))
    return out


def langevin_noise(seed: int, step: int):
    """The Langevin thermostat's Gaussian 3-vector of a noise step (host draw)."""
    g = np.zeros(3)
    L.lib().rbmd_langevin_noise_host(int(seed), int(step), _p(g))
    return g


def fcc(nc: int, rho: float):
    n = 4 * nc ** 3
    x, y, z = (np.zeros(n) for _ in range(3))
    Lb = C.c_double()
    rc = L.lib().rbmd_sys_fcc(nc, rho, _p(x), _p(y), _p(z), C.byref(Lb))
    assert rc == 0
    return x, y, z, Lb.value


def random_ions(n: int, rho: float, dmin: float = 0.8, seed: int = DEFAULT_SEED):
    Lb = (n / rho) ** (1.0 / 3.0)
    x, y, z = (np.zeros(n) for _ in range(3))
    rc = L.lib().rbmd_sys_random(n, Lb, dmin, seed, _p(x), _p(y), _p(z))
    if rc != 0:
        raise RuntimeError("random packing failed")
    q = np.zeros(n)
    assert L.lib().rbmd_sys_charges(n, seed + 1, _p(q)) == 0
    return x, y, z, q, Lb


def velocities(n: int, T: float, seed: int = DEFAULT_SEED, mass=None):
    vx, vy, vz = (np.zeros(n) for _ in range(3))
    mp = None if mass is None else _p(np.ascontiguousarray(mass, np.float64))
    rc = L.lib().rbmd_sys_velocities(n, T, mp, seed + 2, _p(vx), _p(vy), _p(vz))
    assert rc == 0
    return vx, vy, vz


def sample_kvectors_host(alpha: float, box_len, P: int, seed: int, step: int):
    """The device sampler's draws computed on the host: (m[P,3] int32, S)."""
    Lb = np.ascontiguousarray(box_len, np.float64)
    m = np.zeros((P, 3), np.int32)
    S = C.c_double()
    rc = L.lib().rbmd_sample_kvectors_host(alpha, _p(Lb), P, seed, step, _p(m), C.byref(S))
    assert rc == 0
    return m, S.value


def replica_sizes(count: int, n_min: int = 2048, n_max: int = 65536,
                  seed: int = DEFAULT_SEED):
    """Log-uniform replica sizes: N_r = floor(n_min (n_max/n_min)^u_r), u_r in [0, 1)
    the 53-bit uniform of Philox4x32-10 at counter (r, 0x52455053, 0, 0), key = seed."""
    key = (seed & 0xFFFFFFFF, seed >> 32)
    out = []
    for r in range(count):
        w = philox4x32((r, 0x52455053, 0, 0), key)
        u = ((int(w[0]) >> 5) * 67108864.0 + (int(w[1]) >> 6)) / 9007199254740992.0
        out.append(int(np.floor(n_min * (n_max / n_min) ** u)))
    return out
