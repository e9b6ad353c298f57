"""Host-side mirror of the reference Force interface over the HIP C ABI.

`LJCutCoulKspaceForce` keeps the shape of rbmd-gpu's class of the same name
(rbmd-gpu src/force/include/lj_cut_coul_kspace_force.h, Init/Execute of
src/force/include/force.h:8-25): construction takes the configuration the
reference reads in Init() (lj_cut_coul_kspace_force.cpp:37-71: cut_off,
coulomb alpha, coulomb_sample_num, ...), `Init()` binds the per-atom charges
and types, and `Execute()` evaluates LJ + real-space Coulomb + k-space forces
for the current positions.  Differences from the reference, by design:
positions are FP64 SoA device tensors passed to Execute (the reference reads
them from a DataManager singleton), errors raise RuntimeError (the reference
prints and continues, or exit()s; rbmd_define.h:129-150), and atom order is
preserved (the reference permutes all atom arrays into cell order).

Device memory and streams come from torch; every computation runs in
librbmd_force.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L


class RBMDError(RuntimeError):
    pass


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, torch.cuda.Stream):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


@dataclass
class ForceConfig:
    """Configuration of one charged-LJ force evaluator (reference JSON key names in comments)."""
    box_len: tuple
    box_lo: tuple = (0.0, 0.0, 0.0)
    cut_off: float = 2.5          # "cut_off" (LJ)
    coulomb_cut: float = 5.0      # real-space Coulomb cutoff
    skin: float = 0.3             # Verlet skin
    alpha: float = 0.5            # "alpha"
    kspace: str = "RBE"           # "RBE" | "EWALD" | "NONE"  ("coulomb.type")
    coulomb_sample_num: int = 100  # "coulomb_sample_num" = P
    precision: int = 32
    eps: tuple = (1.0,)
    sigma: tuple = (1.0,)
    qqr2e: float = 1.0
    ewald_tol: float = 1e-16
    kmax: int = 2                 # "coulomb.kmax": lattice cube of the thermo far energy
    neighbor_type: str = "VERLET-LIST"  # "neighbor.type": "VERLET-LIST" | "RBL"
    r_core: float = 0.0           # "neighbor.r_core" (RBL core radius)
    rbl_pice: int = 1             # RBL shell stride (rbl_pice_num of "neighbor_sample_num")
    rbl_seed: int = 0             # RBL shell selection stream
    half_shell: bool = False      # True: FP32 pair pass over half-shell rows (each pair once)
    n_max: int = 0
    list_capacity: int = 0        # Verlet row capacity; 0 = automatic
    tile_stage: int = 0           # LDS stage of a cell tile (atoms); 0 = automatic
    extra: dict = field(default_factory=dict)

    @property
    def r_list(self) -> float:
        return max(self.cut_off, self.coulomb_cut) + self.skin


class LJCutCoulKspaceForce:
    _KS = {"NONE": L.KSPACE_NONE, "RBE": L.KSPACE_RBE, "EWALD": L.KSPACE_EWALD}

    def __init__(self, cfg: ForceConfig, _borrowed=None):
        self.cfg = cfg
        self.lib = L.lib()
        self.n = 0
        self.fdtype = torch.float32 if cfg.precision == 32 else torch.float64
        if _borrowed is not None:
            # a replica handle owned by a ReplicaBatch
            self.h = _borrowed
            self._owned = False
            return
        self._owned = True
        p, self._keep = self.params(cfg)
        h = C.c_void_p()
        rc = self.lib.rbmd_force_create(C.byref(p), int(cfg.n_max), C.byref(h))
        if rc != 0:
            raise RBMDError(f"rbmd_force_create -> {rc}: "
                            f"{self.lib.rbmd_force_last_error(None).decode()}")
        self.h = h

    @classmethod
    def params(cls, cfg: ForceConfig):
        """(rbmd_force_params, arrays it points to) for a configuration."""
        p = L.ForceParams()
        lo = [float(v) for v in cfg.box_lo]
        ln = [float(v) for v in cfg.box_len]
        p.box_lo = (C.c_double * 3)(*lo)
        p.box_len = (C.c_double * 3)(*ln)
        p.rc_lj = cfg.cut_off
        p.rc_coul = cfg.coulomb_cut
        p.skin = cfg.skin
        p.alpha = cfg.alpha
        p.qqr2e = cfg.qqr2e
        eps = (C.c_double * len(cfg.eps))(*cfg.eps)
        sig = (C.c_double * len(cfg.sigma))(*cfg.sigma)
        p.ntypes = len(cfg.eps)
        p.eps = C.cast(eps, C.POINTER(C.c_double))
        p.sigma = C.cast(sig, C.POINTER(C.c_double))
        p.precision = cfg.precision
        p.kspace = cls._KS[cfg.kspace.upper()]
        p.p_max = cfg.coulomb_sample_num if cfg.kspace.upper() == "RBE" else 0
        p.ewald_tol = cfg.ewald_tol
        p.list_capacity = cfg.list_capacity
        p.tile_stage = cfg.tile_stage
        rbl = cfg.neighbor_type.upper() == "RBL"
        p.neighbor_type = L.NEIGHBOR_RBL if rbl else L.NEIGHBOR_VERLET
        p.r_core = float(cfg.r_core)
        p.rbl_pice = int(cfg.rbl_pice)
        p.rbl_seed = int(cfg.rbl_seed)
        p.half_shell = 1 if cfg.half_shell else 0
        return p, (eps, sig)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.rbmd_force_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.rbmd_force_last_error(self.h).decode()
            raise RBMDError(f"{what} -> {rc}: {msg}")

    # ------------------------------------------------------------ reference interface
    def Init(self, charge: torch.Tensor, atom_type: torch.Tensor | None = None, stream=None):
        """Bind charges (FP64 device tensor) and types (int32 device tensor or None)."""
        assert charge.is_cuda and charge.dtype == torch.float64 and charge.is_contiguous()
        if atom_type is not None:
            assert atom_type.is_cuda and atom_type.dtype == torch.int32
        self.n = charge.numel()
        self.charge = charge  # kept for the handle-free k-space step (thermo far energy)
        self._check(self.lib.rbmd_force_set_atoms(self.h, self.n, _ptr(charge), _ptr(atom_type),
                                                  _stream(stream)), "set_atoms")

    def Execute(self, x, y, z, fx, fy, fz, energy_virial=False, stream=None):
        """Forces of FP64 positions x, y, z into fx, fy, fz (float32 or float64 by precision)."""
        for t in (x, y, z):
            assert t.is_cuda and t.dtype == torch.float64 and t.numel() == self.n
        for t in (fx, fy, fz):
            assert t.is_cuda and t.dtype == self.fdtype and t.numel() == self.n
        self._check(self.lib.rbmd_force_compute(self.h, _ptr(x), _ptr(y), _ptr(z), _ptr(fx),
                                                _ptr(fy), _ptr(fz),
                                                L.RBMD_EV if energy_virial else 0,
                                                _stream(stream)), "compute")

    # ------------------------------------------------------------ k-space
    def sample_kvectors(self, P: int, seed: int, step: int, stream=None):
        self._check(self.lib.rbmd_force_sample_kvectors(self.h, P, seed, step, _stream(stream)),
                    "sample_kvectors")

    def set_kvectors(self, m: torch.Tensor, w: torch.Tensor | None = None, stream=None):
        assert m.is_cuda and m.dtype == torch.int32 and m.numel() % 3 == 0
        P = m.numel() // 3
        if w is not None:
            assert w.is_cuda and w.dtype == torch.float64 and w.numel() == P
        self._check(self.lib.rbmd_force_set_kvectors(self.h, P, _ptr(m), _ptr(w),
                                                     _stream(stream)), "set_kvectors")

    def kspace(self):
        """(m[P,3], w[P], re[P], im[P]) as numpy arrays, after the last compute."""
        P = C.c_int32()
        self._check(self.lib.rbmd_force_get_kspace(self.h, C.byref(P), None, None, None, None),
                    "get_kspace")
        n = P.value
        m = np.zeros((n, 3), np.int32)
        w = np.zeros(n)
        re = np.zeros(n)
        im = np.zeros(n)
        self._check(self.lib.rbmd_force_get_kspace(
            self.h, C.byref(P), m.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
            re.ctypes.data_as(C.c_void_p), im.ctypes.data_as(C.c_void_p)), "get_kspace")
        return m, w, re, im

    # ------------------------------------------------------------ outputs
    def energy_virial(self, stream=None) -> dict:
        ev = L.EnergyVirial()
        self._check(self.lib.rbmd_force_get_energy_virial(self.h, C.byref(ev), _stream(stream)),
                    "get_energy_virial")
        return {"e_lj": ev.e_lj, "e_real": ev.e_real, "e_self": ev.e_self,
                "e_kspace": ev.e_kspace, "virial": np.array(list(ev.virial))}

    def neighbors(self):
        """Full Verlet list as CSR (offsets[n+1], idx[E]) in original atom ids."""
        offs = np.zeros(self.n + 1, np.int64)
        cnt = C.c_int64()
        self._check(self.lib.rbmd_force_get_neighbors(self.h, offs.ctypes.data_as(C.c_void_p),
                                                      None, 0, C.byref(cnt)), "get_neighbors")
        idx = np.zeros(max(cnt.value, 1), np.int32)
        self._check(self.lib.rbmd_force_get_neighbors(self.h, offs.ctypes.data_as(C.c_void_p),
                                                      idx.ctypes.data_as(C.c_void_p), idx.size,
                                                      C.byref(cnt)), "get_neighbors")
        return offs, idx[:cnt.value]

    def stats(self) -> dict:
        out = np.zeros(14, np.int64)
        self._check(self.lib.rbmd_force_get_stats(self.h, out.ctypes.data_as(C.c_void_p)),
                    "get_stats")
        keys = ["builds", "entries", "cap", "overflow_rows", "cells", "inner_entries",
                "ewald_k", "n", "big_tiles", "tile_stage", "largest_tile", "split",
                "rho_atoms_per_workgroup", "tuned_cost_us"]
        return dict(zip(keys, (int(v) for v in out)))

    def autotune(self, x, y, z, fx, fy, fz, reps=5, stream=None):
        """Pick tile shapes (workgroups per tile, LDS stage, rho(k) workgroup size) by
        timing candidates on this configuration; returns the chosen shapes."""
        self._check(self.lib.rbmd_force_autotune(self.h, _ptr(x), _ptr(y), _ptr(z), _ptr(fx),
                                                 _ptr(fy), _ptr(fz), reps, _stream(stream)),
                    "autotune")
        st = self.stats()
        return {k: st[k] for k in ("split", "tile_stage", "rho_atoms_per_workgroup",
                                   "tuned_cost_us")}

    def set_timing(self, enable: bool):
        self._check(self.lib.rbmd_force_set_timing(self.h, 1 if enable else 0), "set_timing")

    def timing(self, name: str):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(self.lib.rbmd_force_get_timing(self.h, name.encode(), C.byref(ms),
                                                   C.byref(cnt)), "get_timing")
        return ms.value, cnt.value

    def time_pair(self, x, y, z, fx, fy, fz, reps=20, stream=None):
        """Mean device ms of the pair pass alone at (x, y, z): reps launches back to
        back between one pair of HIP events; fx, fy, fz end holding the forces."""
        for t in (x, y, z):
            assert t.is_cuda and t.dtype == torch.float64 and t.numel() == self.n
        for t in (fx, fy, fz):
            assert t.is_cuda and t.dtype == self.fdtype and t.numel() == self.n
        ms = C.c_double()
        self._check(self.lib.rbmd_force_time_pair(self.h, _ptr(x), _ptr(y), _ptr(z), _ptr(fx),
                                                  _ptr(fy), _ptr(fz), reps, C.byref(ms),
                                                  _stream(stream)), "time_pair")
        return ms.value

    # ------------------------------------------------------------ velocity Verlet
    _THERMO = {"NONE": L.THERMO_NONE, "BERENDSEN": L.THERMO_BERENDSEN,
               "TEMP_RESCALE": L.THERMO_RESCALE, "NOSE_HOOVER": L.THERMO_NOSE_HOOVER,
               "LANGEVIN": L.THERMO_LANGEVIN, "MEASURE": L.THERMO_MEASURE}

    def set_thermostat(self, temp_ctrl_type="NONE", temperature=(1.0, 1.0, 0.5), dof=0.0,
                       kb=1.0, xi=0.0, gamma=100.0, kbt=1.0, seed=0, noise_step=0):
        """NVT thermostat of md_run, keyed as the reference JSON ("temp_ctrl_type",
        "temperature" = [T_target, T_end, Tdamp]; ExecutionNVT::InitParameters,
        run/ExecutionNVT.cpp:840-841).  dof 0 means 3n - 3."""
        t = L.Thermostat(self._THERMO[temp_ctrl_type], float(temperature[0]),
                         float(temperature[2]), float(dof), float(kb), float(xi), float(gamma),
                         float(kbt), int(seed), int(noise_step))
        self._check(self.lib.rbmd_md_set_thermostat(self.h, C.byref(t)), "set_thermostat")

    def thermostat_state(self) -> dict:
        """Temperature of the last step (before its rescale), Nose-Hoover xi, next noise step."""
        T, xi, ns = C.c_double(), C.c_double(), C.c_uint64()
        self._check(self.lib.rbmd_md_get_thermostat(self.h, C.byref(T), C.byref(xi),
                                                    C.byref(ns)), "get_thermostat")
        return {"T": T.value, "xi": xi.value, "noise_step": ns.value}

    def set_temperature_log(self, log: torch.Tensor | None):
        """FP64 device tensor receiving the temperature of each step of later runs."""
        if log is not None:
            assert log.dtype == torch.float64 and log.is_cuda and log.is_contiguous()
        self._temp_log = log  # keep the memory alive while the handle may write it
        n = 0 if log is None else log.numel()
        self._check(self.lib.rbmd_md_set_temperature_log(self.h, _ptr(log), n),
                    "set_temperature_log")

    def md_run(self, x, y, z, vx, vy, vz, fx, fy, fz, dt, nsteps, seed=0, step0=0, img=None,
               mass=None, use_graph=True, stream=None):
        """nsteps of kick-drift-force-kick (ExecutionNVT::Solve order) updating all arrays."""
        st = _stream(stream)
        self._check(self.lib.rbmd_md_run(self.h, _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy),
                                         _ptr(vz), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(img),
                                         _ptr(mass), dt, nsteps, seed, step0,
                                         1 if use_graph else 0, st), "md_run")


def rbl_pice_num(n: int, box_len, cut_off: float, r_core: float, neighbor_sample_num: float):
    """The random batch list's shell stride as the reference derives it from
    "neighbor_sample_num" (ExecutionMD::ComputeRBLNearForce,
    rbmd/framework/run/ExecutionMD.cpp:476-484): density rho = n / V (FP32 `Real`),
    coeff = round(1 + 0.05 / rho - 0.05), counts c(r) = trunc(coeff rho ceil(4/3 pi r^3) + 1),
    pice = ceil(1 / (neighbor_sample_num / (c(cut_off) - c(r_core))))."""
    V = float(box_len[0]) * float(box_len[1]) * float(box_len[2])
    rho = float(np.float32(n / V))
    coeff = float(np.floor(1.0 + (0.05 / rho - 0.05) + 0.5))

    def count(r):
        return int(coeff * rho * math.ceil(4.0 / 3.0 * float(np.float32(math.pi)) * r ** 3) + 1)

    rate = float(neighbor_sample_num) / float(count(cut_off) - count(r_core))
    return max(1, int(math.ceil(1.0 / rate)))


def kspace_compute(box_len, alpha, x, y, z, q, m, w=None, fx=None, fy=None, fz=None,
                   box_lo=(0.0, 0.0, 0.0), qqr2e=1.0, precision=64, stream=None,
                   noalloc=False):
    """The k-space step of a given k-list without a force handle (rbmd_kspace_compute):
    returns (rho[P, 2] FP64, ev[7] = E_k, W_xx, W_yy, W_zz, W_xy, W_xz, W_yz), and adds
    the k-space force to fx, fy, fz when given (scratch arrays otherwise).
    m: int32 device tensor (3P); w: FP64 device tensor (P) or None (S/P).
    noalloc: run rbmd_kspace_compute_noalloc (the library's fixed per-device scratch,
    P <= 8192, no synchronisation; the output tensors here are still allocated by torch)."""
    lib = L.lib()
    n, P = x.numel(), m.numel() // 3
    dt = torch.float32 if precision == 32 else torch.float64
    if fx is None:
        fx, fy, fz = (torch.zeros(n, device=x.device, dtype=dt) for _ in range(3))
    rho = torch.zeros(2 * P, device=x.device, dtype=torch.float64)
    ev = torch.zeros(7, device=x.device, dtype=torch.float64)
    prm = L.KspaceParams((C.c_double * 3)(*[float(v) for v in box_lo]),
                         (C.c_double * 3)(*[float(v) for v in box_len]), float(alpha),
                         float(qqr2e), int(precision))
    fn = lib.rbmd_kspace_compute_noalloc if noalloc else lib.rbmd_kspace_compute
    rc = fn(C.byref(prm), n, _ptr(x), _ptr(y), _ptr(z), _ptr(q), P, _ptr(m), _ptr(w), _ptr(fx),
            _ptr(fy), _ptr(fz), _ptr(rho), _ptr(ev), _stream(stream))
    if rc != 0:
        name = "rbmd_kspace_compute_noalloc" if noalloc else "rbmd_kspace_compute"
        raise RBMDError(f"{name} -> {rc}: {lib.rbmd_kspace_last_error().decode()}")
    return rho.view(P, 2), ev
