"""Thermo output rows of the reference (ThermoOutput, rbmd/framework/output/ThermoOutput.cpp).

A row holds per-atom energies (:152-366): VanderWaalsEnergy = E_LJ / N,
NearCoulombicEnergy = E_real / N, FarCoulombicEnergy = (E_k + E_self) / N,
PotentialEnergy = their sum, KinticEnergy = sum m v^2 / N / 2 with the sum taken after
the closing kick of the last step (ExecutionNVT::ComputeTempe), TotalEnergy, and
Residual = the change of PotentialEnergy since the previous row (:410-421).

The LJ and near energies come from one force pass with energies and virial on scratch
force arrays (the reference also runs a separate energy pass per row); the kinetic
energy from the temperature the library measured in the last MD step, so the handle
needs a thermostat (RBMD_THERMO_MEASURE measures without touching the velocities).
The far energy is, as in the reference whatever sampler drives the dynamics, the
deterministic Ewald sum over the lattice cube |m_d| <= Kmax, m != 0 (:218-252):
E_k = (2 pi / V) sum exp(-k^2 / 4 alpha) |rho(k)|^2 / k^2, computed on the device by
the handle-free k-space step (rbmd_kspace_compute) with weights exp(-k^2 / 4 alpha),
plus the self energy -sqrt(alpha / pi) sum q^2.
"""
from __future__ import annotations

import numpy as np
import torch

from .force import kspace_compute

COLUMNS = ("Step", "Time", "VanderWaalsEnergy", "NearCoulombicEnergy", "FarCoulombicEnergy",
           "Residual", "KinticEnergy", "PotentialEnergy", "TotalEnergy", "Temperature")


def kmax_cube(box_len, alpha: float, kmax: int):
    """Lattice vectors of the cube |m_d| <= kmax without m = 0 (the reference's loop
    order, x outermost) and their weights exp(-k^2 / 4 alpha)."""
    r = np.arange(-kmax, kmax + 1)
    a, b, c = np.meshgrid(r, r, r, indexing="ij")
    m = np.stack([a.ravel(), b.ravel(), c.ravel()], 1)
    m = m[(m != 0).any(1)]
    k = 2.0 * np.pi * m / np.asarray(box_len, dtype=np.float64)
    return m.astype(np.int32), np.exp(-(k * k).sum(1) / (4.0 * alpha))


def far_kspace_energy(f, X, Y, Z, kmax: int | None = None) -> float:
    """E_k of the Kmax cube at the current positions (qqr2e applied), FP64."""
    cfg = f.cfg
    kmax = cfg.kmax if kmax is None else int(kmax)
    m, w = kmax_cube(cfg.box_len, cfg.alpha, kmax)
    dev = X.device
    _, ev = kspace_compute(cfg.box_len, cfg.alpha, X, Y, Z, f.charge,
                           torch.as_tensor(m.reshape(-1)).to(dev),
                           torch.as_tensor(w).to(dev, torch.float64), box_lo=cfg.box_lo,
                           qqr2e=cfg.qqr2e, precision=64)
    return float(ev[0])


def thermo_row(f, X, Y, Z, step: int, dt: float, previous: dict | None = None, dof=None,
               kb: float = 1.0, kmax: int | None = None) -> dict:
    """One row at the current positions; previous is the row before (for Residual);
    kmax overrides the configuration's Kmax ("coulomb.kmax")."""
    n = X.numel()
    F = [torch.empty(n, device=X.device, dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F, energy_virial=True)
    ev = f.energy_virial()
    lj, near = ev["e_lj"] / n, ev["e_real"] / n
    far = 0.0
    if f.cfg.kspace.upper() != "NONE":
        far = (far_kspace_energy(f, X, Y, Z, kmax) + ev["e_self"]) / n
    pe = lj + near + far
    T = f.thermostat_state()["T"]
    dof = 3.0 * n - 3.0 if dof is None else float(dof)
    ke = 0.5 * T * dof * kb / n
    return {"Step": int(step), "Time": step * dt, "VanderWaalsEnergy": lj,
            "NearCoulombicEnergy": near, "FarCoulombicEnergy": far,
            "Residual": 0.0 if previous is None else pe - previous["PotentialEnergy"],
            "KinticEnergy": ke, "PotentialEnergy": pe, "TotalEnergy": pe + ke,
            "Temperature": T}


def format_rows(rows) -> str:
    """Header line and one space-separated line per row, in the reference's column order."""
    out = [" ".join(COLUMNS)]
    for r in rows:
        out.append(" ".join(repr(r[c]) for c in COLUMNS))
    return "\n".join(out) + "\n"
