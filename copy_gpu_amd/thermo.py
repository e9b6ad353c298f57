"""Thermo output rows of the reference (ThermoOutput, rbmd/framework/output/ThermoOutput.cpp).

A row holds per-atom energies (:152-366): VanderWaalsEnergy = E_LJ / N,
NearCoulombicEnergy = E_real / N, FarCoulombicEnergy = (E_k + E_self) / N,
PotentialEnergy = their sum, KinticEnergy = sum m v^2 / N / 2 with the sum taken after
the closing kick of the last step (ExecutionNVT::ComputeTempe), TotalEnergy, and
Residual = the change of PotentialEnergy since the previous row (:410-421).

The energies come from one force pass with energies and virial on scratch force arrays
(the reference also runs a separate energy pass per row); the kinetic energy from the
temperature the library measured in the last MD step, so the handle needs a thermostat
(RBMD_THERMO_MEASURE measures without touching the velocities).  The far-field energy
is that of the handle's k-space: the RBE estimate for an RBE handle (unbiased), the
lattice sum for an Ewald handle; the reference sums the Ewald lattice up to Kmax
(:218-252).
"""
from __future__ import annotations

import torch

COLUMNS = ("Step", "Time", "VanderWaalsEnergy", "NearCoulombicEnergy", "FarCoulombicEnergy",
           "Residual", "KinticEnergy", "PotentialEnergy", "TotalEnergy", "Temperature")


def thermo_row(f, X, Y, Z, step: int, dt: float, previous: dict | None = None, dof=None,
               kb: float = 1.0) -> dict:
    """One row at the current positions; previous is the row before (for Residual)."""
    n = X.numel()
    F = [torch.empty(n, device=X.device, dtype=f.fdtype) for _ in range(3)]
    f.Execute(X, Y, Z, *F, energy_virial=True)
    ev = f.energy_virial()
    lj, near = ev["e_lj"] / n, ev["e_real"] / n
    far = (ev["e_kspace"] + ev["e_self"]) / n
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
