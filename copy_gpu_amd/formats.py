"""Input and output formats of the reference, over the library's C ABI.

- `read_data` reads a LAMMPS data file as the reference's "read_data" does
  (rbmd-gpu src/reader/src/atomic_reader.cpp:24-58, 135-230, 322-350): header
  counts and box, Masses, Pair Coeffs, Atoms of style atomic / charge / full,
  Velocities.  The atom with id k lands at index k - 1 and types are 0-based.
- `write_dump` appends trajectory frames in the layout of TrajectoryOutput
  (rbmd-parallel/framework/output/TrajectoryOutput.cpp:21-44).
- `load_config` maps the reference's JSON input (res/json_files/*/rbmd.json; "//"
  comments and trailing commas allowed, as the reference's own files have them)
  onto a ForceConfig and the run settings of the VV loop and thermostat.

The parsing and writing run in librbmd_force.so; this module only allocates.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import re

import numpy as np

from . import _lib as L
from .force import ForceConfig, RBMDError, rbl_pice_num


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what):
    if rc != L.RBMD_OK:
        raise RBMDError(f"{what} -> {rc}: {L.lib().rbmd_io_last_error().decode()}")


def read_data(path: str, atom_style: str = "charge") -> dict:
    """Arrays of a LAMMPS data file, indexed by atom id - 1 (HOST numpy)."""
    info = L.DataInfo()
    bpath = os.fsencode(path)
    _check(L.lib().rbmd_read_data_info(bpath, C.byref(info)), "read_data_info")
    n, nt = int(info.natoms), int(info.ntypes)
    out = {
        "n": n, "ntypes": nt,
        "lo": np.array(info.lo[:]), "hi": np.array(info.hi[:]),
        "type": np.zeros(n, np.int32), "q": np.zeros(n), "x": np.zeros(n), "y": np.zeros(n),
        "z": np.zeros(n), "img": np.zeros(3 * n, np.int32),
        "mass": np.ones(nt), "eps": np.ones(nt), "sigma": np.ones(nt),
        "has_masses": bool(info.has_masses), "has_pair_coeffs": bool(info.has_pair_coeffs),
        "has_velocities": bool(info.has_velocities),
    }
    v = [np.zeros(n) for _ in range(3)]
    _check(L.lib().rbmd_read_data(
        bpath, atom_style.encode(), n, _p(out["type"]), _p(out["q"]), _p(out["x"]),
        _p(out["y"]), _p(out["z"]), _p(out["img"]), _p(v[0]), _p(v[1]), _p(v[2]),
        _p(out["mass"]), _p(out["eps"]), _p(out["sigma"])), "read_data")
    out["vx"], out["vy"], out["vz"] = v
    return out


def write_dump(path: str, step: int, lo, hi, atom_type, x, y, z, append: bool = True):
    """One frame: ITEM: TIMESTEP / NUMBER OF ATOMS / BOX BOUNDS pp pp pp / ATOMS id type x y z."""
    t = np.ascontiguousarray(atom_type, np.int32)
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    lo, hi = (np.ascontiguousarray(a, np.float64) for a in (lo, hi))
    _check(L.lib().rbmd_write_dump(os.fsencode(path), 1 if append else 0, int(step), t.size,
                                   _p(lo), _p(hi), _p(t), _p(x), _p(y), _p(z)), "write_dump")


def _strip_json(text: str) -> str:
    out, i, n, in_str = [], 0, len(text), False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 1
            elif c == '"':
                in_str = False
        elif c == '"':
            in_str = True
            out.append(c)
        elif text.startswith("//", i):
            while i < n and text[i] != "\n":
                i += 1
            continue
        else:
            out.append(c)
        i += 1
    s = "".join(out).lstrip("﻿")
    return re.sub(r",(\s*[}\]])", r"\1", s)  # trailing commas


def load_config(path: str, n_max: int = 0, skin: float = 0.3, coulomb_cut: float | None = None,
                precision: int = 32) -> tuple[ForceConfig, dict, dict]:
    """(ForceConfig, run settings, raw JSON) from a reference JSON input file.

    The reference has one cutoff ("neighbor"."cut_off") for LJ and real-space
    Coulomb; coulomb_cut overrides the second.  A "read_data" initial
    configuration is read from the file it names (relative to the JSON), and its
    box, masses and Pair Coeffs fill the configuration."""
    with open(path, encoding="utf-8-sig") as f:
        cfg = json.loads(_strip_json(f.read()))
    hp = cfg.get("hyper_parameters", {})
    nb, cl = hp.get("neighbor", {}), hp.get("coulomb", {})
    ex = cfg.get("execution", {})
    rc = float(nb.get("cut_off", 2.5))
    kspace = cl.get("type", "NONE").upper() if cl else "NONE"
    run = {
        "dt": float(ex.get("timestep", 0.002)),
        "num_steps": int(ex.get("num_steps", 0)),
        "ensemble": ex.get("ensemble", "NVE"),
        "temp_ctrl_type": ex.get("temp_ctrl_type", "NONE") if ex.get("ensemble") == "NVT"
        else "NONE",
        "temperature": tuple(float(t) for t in ex.get("temperature", (1.0, 1.0, 0.5))),
        "neighbor_type": nb.get("type", "VERLET-LIST"),
        "trajectory_interval": int(cfg.get("outputs", {}).get("trajectory_out", {})
                                   .get("interval", 0)),
    }
    init = cfg.get("init_configuration", {})
    box_lo, box_len, eps, sigma = (0.0, 0.0, 0.0), None, (1.0,), (1.0,)
    if "read_data" in init:
        rd = init["read_data"]
        data = read_data(os.path.join(os.path.dirname(os.path.abspath(path)), rd["file"]),
                         rd.get("atom_style", "charge"))
        run["data"] = data
        box_lo = tuple(float(v) for v in data["lo"])
        box_len = tuple(float(v) for v in data["hi"] - data["lo"])
        eps, sigma = tuple(data["eps"]), tuple(data["sigma"])
        n_max = max(n_max, data["n"])
    elif "inbuild" in init:
        ib = init["inbuild"]
        box_lo = tuple(float(ib[k][0]) for k in ("x_range", "y_range", "z_range"))
        box_len = tuple(float(ib[k][1] - ib[k][0]) for k in ("x_range", "y_range", "z_range"))
        run["inbuild_dims"] = tuple(int(d) for d in ib.get("dims", (10, 10, 10)))
    if box_len is None:
        raise RBMDError("load_config: no init_configuration (read_data or inbuild)")
    rbl = {}
    if str(nb.get("type", "")).upper() == "RBL":
        # random batch list: core radius and the shell stride the reference derives
        # from neighbor_sample_num at the system's density
        n_atoms = run["data"]["n"] if "data" in run else int(np.prod(run["inbuild_dims"]))
        r_core = float(nb.get("r_core", rc / 2))
        rbl = {"neighbor_type": "RBL", "r_core": r_core,
               "rbl_pice": rbl_pice_num(n_atoms, box_len, rc, r_core,
                                        float(nb.get("neighbor_sample_num", 50)))}
        run["neighbor_sample_num"] = float(nb.get("neighbor_sample_num", 50))
    fc = ForceConfig(box_len=box_len, box_lo=box_lo, cut_off=rc,
                     coulomb_cut=rc if coulomb_cut is None else float(coulomb_cut), skin=skin,
                     alpha=float(cl.get("alpha", 0.5)), kspace=kspace,
                     coulomb_sample_num=int(cl.get("coulomb_sample_num", 100)),
                     kmax=int(cl.get("kmax", 2)), **rbl,
                     precision=precision, eps=eps, sigma=sigma, n_max=n_max)
    return fc, run, cfg


def inbuild_system(run: dict, box_lo, box_len):
    """The reference's in-built initial configuration ("init_configuration.inbuild"):
    one atom at the centroid of every cell of a dims grid over the box, x fastest
    (MeshFreeCondition::InitPosition, rbmd/framework/InitCondition/MeshFreeCondition.cpp:58-80,
    the cell centroids of a uniform data set), with charge -1 on the first half of the
    atoms and +1 on the rest (ExecutionMD::InitialCondition,
    rbmd/framework/run/ExecutionMD.cpp:155-168).  Returns (x, y, z, q)."""
    dims = run["inbuild_dims"]
    ax = [box_lo[d] + (np.arange(dims[d]) + 0.5) * (box_len[d] / dims[d]) for d in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    n = x.size
    q = np.full(n, -1.0)
    q[n // 2:] = 1.0
    return x.ravel().copy(), y.ravel().copy(), z.ravel().copy(), q
