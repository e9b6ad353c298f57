"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

numpy bindings of liboracle.so (refcpu.c: FP64 restatement of the reference
RBMD CPU path; faithful.c: reference-faithful FP32 step for the CPU baseline).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    r = subprocess.run(["make", "-s", "-B", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        d, i64, i32, P = C.c_double, C.c_int64, C.c_int32, C.c_void_p
        sig = {
            "orc_wrap": (None, [i64, P, P, P, P, P, P]),
            "orc_neighbors_brute": (i64, [i64, P, P, P, P, d, P, P, i64]),
            "orc_neighbors_cells": (i64, [i64, P, P, P, P, P, d, P, P, i64]),
            "orc_gnear": (d, [d, d]),
            "orc_pair_forces": (None, [i64, P, P, P, P, P, P, P, P, P, d, d, d, d, P, P, P, P]),
            "orc_rbe_S": (d, [d, P, P]),
            "orc_rho": (None, [i64, P, P, P, P, P, i32, P, P, P]),
            "orc_kspace_forces": (None, [i64, P, P, P, P, P, d, d, i32, P, P, P, P, P, P, P, P]),
            "orc_ewald_kvectors": (i64, [P, d, P, P, P]),
            "orc_self_energy": (d, [i64, P, d, d]),
            "orc_kick": (None, [i64, P, P, P, P, P, P, P, d, d]),
            "orc_drift": (None, [i64, P, P, P, P, P, P, d, P, P, P]),
            "orc_philox4x32": (None, [P, P, P]),
            "orc_rbl_select": (C.c_int, [C.c_uint64, C.c_uint32, i32, i32, C.c_uint32]),
            "orc_rbl_forces": (None, [i64, P, P, P, P, P, P, P, P, P, d, d, d, d, d, i32,
                                      C.c_uint64, C.c_uint32, P, P, P, P]),
            "orc_neighbors_images": (i64, [i64, P, P, P, P, d, P, P, i64]),
            "orc_faithful_run": (C.c_int, [i64, P, P, P, P, P, P, d, d, d, i32, P, i32, d, i32]),
            "orc_faithful_run_f64": (C.c_int, [i64, P, P, P, P, P, P, d, d, d, i32, P, i32, d,
                                               i32]),
            "orc_num_threads": (C.c_int, []),
            "orc_temperature": (d, [i64, P, P, P, P, d, d, d]),
            "orc_scale_velocities": (None, [i64, P, P, P, d]),
            "orc_rescale_coeff": (d, [d, d]),
            "orc_berendsen_coeff": (d, [d, d, d, d]),
            "orc_nose_hoover_kick": (None, [i64, P, P, P, P, P, P, P, d, d, d]),
            "orc_nose_hoover_xi": (d, [d, d, d, d]),
            "orc_langevin_force": (None, [i64, P, P, P, P, P, P, P, d, d, d, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(_lib, k)
            f.restype = r
            f.argtypes = a
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def wrap(x, y, z, lo, Lb):
    x, y, z = _d(x).copy(), _d(y).copy(), _d(z).copy()
    lib().orc_wrap(x.size, _p(x), _p(y), _p(z), _p(_d(lo)), _p(_d(Lb)), None)
    return x, y, z


def _csr(fn, n, *args):
    offs = np.zeros(n + 1, np.int64)
    cap = max(64, n * 64)
    while True:
        idx = np.zeros(cap, np.int32)
        e = fn(*args, _p(offs), _p(idx), cap)
        if e == -1 and fn is lib().orc_neighbors_cells:
            raise ValueError("cell grid needs >= 3 cells per axis")
        if e >= 0:
            return offs, idx[:e]
        cap = -e


def neighbors_brute(x, y, z, Lb, r_list):
    x, y, z = _d(x), _d(y), _d(z)
    return _csr(lib().orc_neighbors_brute, x.size, x.size, _p(x), _p(y), _p(z), _p(_d(Lb)),
                float(r_list))


def neighbors_cells(x, y, z, lo, Lb, r_list):
    x, y, z = _d(x), _d(y), _d(z)
    return _csr(lib().orc_neighbors_cells, x.size, x.size, _p(x), _p(y), _p(z), _p(_d(lo)),
                _p(_d(Lb)), float(r_list))


def gnear(alpha, r):
    return lib().orc_gnear(float(alpha), float(r))


def pair_forces(x, y, z, q, lo, Lb, rc_lj, rc_coul, alpha, qqr2e=1.0, types=None,
                eps=(1.0,), sigma=(1.0,)):
    """FP64 LJ + erfc real-space forces; returns (F[n,3], ev8)."""
    x, y, z, q = _d(x), _d(y), _d(z), _d(q)
    n = x.size
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ev = np.zeros(8)
    t = None if types is None else np.ascontiguousarray(types, np.int32)
    lib().orc_pair_forces(n, _p(x), _p(y), _p(z), _p(q), _p(t), _p(_d(eps)), _p(_d(sigma)),
                          _p(_d(lo)), _p(_d(Lb)), float(rc_lj), float(rc_coul), float(alpha),
                          float(qqr2e), _p(fx), _p(fy), _p(fz), _p(ev))
    return np.stack([fx, fy, fz], 1), ev


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(out))
    return out


def rbl_select(seed, epoch, i, j, pice):
    return bool(lib().orc_rbl_select(int(seed), int(epoch), int(i), int(j), int(pice)))


def rbl_forces(x, y, z, q, lo, Lb, rc_lj, rc_coul, alpha, r_core, pice, seed, epoch,
               qqr2e=1.0, types=None, eps=(1.0,), sigma=(1.0,)):
    """Random batch list near forces (core + selected shell, mean removed) and ev8."""
    x, y, z, q = _d(x), _d(y), _d(z), _d(q)
    n = x.size
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ev = np.zeros(8)
    t = None if types is None else np.ascontiguousarray(types, np.int32)
    lib().orc_rbl_forces(n, _p(x), _p(y), _p(z), _p(q), _p(t), _p(_d(eps)), _p(_d(sigma)),
                         _p(_d(lo)), _p(_d(Lb)), float(rc_lj), float(rc_coul), float(alpha),
                         float(qqr2e), float(r_core), int(pice), int(seed), int(epoch), _p(fx),
                         _p(fy), _p(fz), _p(ev))
    return np.stack([fx, fy, fz], 1), ev


def neighbors_images(x, y, z, Lb, r_list):
    """Every periodic image (shifts k L, k in {-1, 0, 1}^3) of every atom within r_list
    of each wrapped atom, as CSR of atom ids (an id once per image)."""
    x, y, z = _d(x), _d(y), _d(z)
    return _csr(lib().orc_neighbors_images, x.size, x.size, _p(x), _p(y), _p(z), _p(_d(Lb)),
                float(r_list))


def rbe_S(alpha, Lb):
    H = np.zeros(3)
    S = lib().orc_rbe_S(float(alpha), _p(_d(Lb)), _p(H))
    return S, H


def rho(x, y, z, q, Lb, m):
    x, y, z, q = _d(x), _d(y), _d(z), _d(q)
    m = np.ascontiguousarray(m, np.int32).reshape(-1, 3)
    P = m.shape[0]
    re, im = np.zeros(P), np.zeros(P)
    lib().orc_rho(x.size, _p(x), _p(y), _p(z), _p(q), _p(_d(Lb)), P, _p(m), _p(re), _p(im))
    return re, im


def kspace_forces(x, y, z, q, Lb, alpha, m, w, re=None, im=None, qqr2e=1.0):
    """k-space forces F[n,3] and ev7 = [E_k, W_xx, W_yy, W_zz, W_xy, W_xz, W_yz]."""
    x, y, z, q = _d(x), _d(y), _d(z), _d(q)
    m = np.ascontiguousarray(m, np.int32).reshape(-1, 3)
    w = _d(w)
    if re is None:
        re, im = rho(x, y, z, q, Lb, m)
    n = x.size
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ev = np.zeros(7)
    lib().orc_kspace_forces(n, _p(x), _p(y), _p(z), _p(q), _p(_d(Lb)), float(alpha),
                            float(qqr2e), m.shape[0], _p(m), _p(w), _p(_d(re)), _p(_d(im)),
                            _p(fx), _p(fy), _p(fz), _p(ev))
    return np.stack([fx, fy, fz], 1), ev


def ewald_kvectors(Lb, alpha, kmax):
    km = np.ascontiguousarray(kmax, np.int32)
    c = lib().orc_ewald_kvectors(_p(_d(Lb)), float(alpha), _p(km), None, None)
    m = np.zeros((c, 3), np.int32)
    w = np.zeros(c)
    lib().orc_ewald_kvectors(_p(_d(Lb)), float(alpha), _p(km), _p(m), _p(w))
    return m, w


def self_energy(q, alpha, qqr2e=1.0):
    q = _d(q)
    return lib().orc_self_energy(q.size, _p(q), float(alpha), float(qqr2e))


def kick(v, f, dt, mass=None, fmt2v=1.0):
    v = [_d(a).copy() for a in v]
    f = [_d(a) for a in f]
    lib().orc_kick(v[0].size, _p(v[0]), _p(v[1]), _p(v[2]), _p(f[0]), _p(f[1]), _p(f[2]),
                   _p(None if mass is None else _d(mass)), float(dt), float(fmt2v))
    return v


def drift(xyz, v, dt, lo, Lb, img=None):
    p = [_d(a).copy() for a in xyz]
    v = [_d(a) for a in v]
    im = None if img is None else np.ascontiguousarray(img, np.int32).copy()
    lib().orc_drift(p[0].size, _p(p[0]), _p(p[1]), _p(p[2]), _p(v[0]), _p(v[1]), _p(v[2]),
                    float(dt), _p(_d(lo)), _p(_d(Lb)), _p(im))
    return p, im


def temperature(v, mass=None, dof=None, kB=1.0, mvv2e=1.0):
    """Kinetic temperature; dof defaults to 3n - 3 (read-in systems without SHAKE)."""
    v = [_d(a) for a in v]
    n = v[0].size
    dof = 3.0 * n - 3.0 if dof is None else float(dof)
    return lib().orc_temperature(n, _p(v[0]), _p(v[1]), _p(v[2]),
                                 _p(None if mass is None else _d(mass)), float(mvv2e), dof,
                                 float(kB))


def scale_velocities(v, c):
    v = [_d(a).copy() for a in v]
    lib().orc_scale_velocities(v[0].size, _p(v[0]), _p(v[1]), _p(v[2]), float(c))
    return v


def rescale_coeff(T, T0):
    return lib().orc_rescale_coeff(float(T), float(T0))


def berendsen_coeff(T, T0, dt, tdamp):
    return lib().orc_berendsen_coeff(float(T), float(T0), float(dt), float(tdamp))


def nose_hoover_kick(v, f, dt, xi, mass=None, fmt2v=1.0):
    v = [_d(a).copy() for a in v]
    f = [_d(a) for a in f]
    lib().orc_nose_hoover_kick(v[0].size, _p(v[0]), _p(v[1]), _p(v[2]), _p(f[0]), _p(f[1]),
                               _p(f[2]), _p(None if mass is None else _d(mass)), float(dt),
                               float(fmt2v), float(xi))
    return v


def nose_hoover_xi(xi, T, T0, dt):
    return lib().orc_nose_hoover_xi(float(xi), float(T), float(T0), float(dt))


def langevin_force(f, v, g, gamma, kbt, dt, mass=None):
    f = [_d(a).copy() for a in f]
    v = [_d(a) for a in v]
    lib().orc_langevin_force(f[0].size, _p(f[0]), _p(f[1]), _p(f[2]), _p(v[0]), _p(v[1]),
                             _p(v[2]), _p(None if mass is None else _d(mass)), float(gamma),
                             float(kbt), float(dt), _p(_d(g)))
    return f


def faithful_run(pos, vel, force, q, lo, Lb, rc_lj, rc_coul, alpha, P, m_steps, nsteps, dt,
                 nthreads=1):
    """Reference-faithful steps (arrays updated in place; AoS [n,3] float32 for the
    FP32 build, float64 for the FP64 build; q None: no charges)."""
    dt_ = pos.dtype
    assert dt_ in (np.float32, np.float64)
    for a in (pos, vel, force):
        assert a.dtype == dt_ and a.flags.c_contiguous
    qf = None if q is None else np.ascontiguousarray(q, dt_)
    m = np.ascontiguousarray(m_steps if m_steps is not None else np.zeros(3), np.int32)
    fn = lib().orc_faithful_run if dt_ == np.float32 else lib().orc_faithful_run_f64
    return fn(pos.shape[0], _p(pos), _p(vel), _p(force), _p(qf), _p(_d(lo)), _p(_d(Lb)),
              float(rc_lj), float(rc_coul), float(alpha), int(P), _p(m), int(nsteps), float(dt),
              int(nthreads))
