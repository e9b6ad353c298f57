"""Write the golden fixtures under tests/golden/ from data files the reference holds.

Run in a checkout that has the reference mounted (default /root/reference); the
committed outputs are data only, so the tests never read the reference:

  gnear_alpha0.1.txt   rows (r, G, dG/dr) of the reference's tabulated real-space
                       Coulomb kernel G(alpha=0.1, r) = erfc(sqrt(a) r)/r^2 +
                       2 sqrt(a) exp(-a r^2)/(sqrt(pi) r), from
                       rbmd/res/json_files/Gnear_table_leq2_2.txt (r in [0.1, 2)) and
                       Gnear_table_geq2_2.txt (r in [2, 11.97]); columns of the source:
                       index, r, step, G, dG/dr.
  lj_salt_6000.npz     positions, charges, types, image flags, velocities and box of
                       the reference example rbmd/res/json_files/LJ_salt/box_102030/rbmd.data
                       (LAMMPS write_data, 6000 +/-1 LJ ions, box 10 x 20 x 30, atom style
                       charge: id type q x y z ix iy iz; Velocities: id vx vy vz), rows
                       sorted by atom id.
  lj_salt_6000.data.gz that example data file itself, gzip-compressed, as the input of
                       the data-file reader's parity test.
  lj_salt_rbmd.json    the reference's example run input rbmd/res/json_files/LJ_salt/rbmd.json
                       (in-built lattice, random batch list, RBE, Berendsen NVT), byte for
                       byte, as the input of the JSON front end's end-to-end test.
"""
from __future__ import annotations

import gzip
import os
import sys

import numpy as np

REF = os.environ.get("RBMD_REFERENCE", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))


def gnear():
    rows = []
    for name in ("Gnear_table_leq2_2.txt", "Gnear_table_geq2_2.txt"):
        path = os.path.join(REF, "rbmd", "res", "json_files", name)
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) != 5:
                    continue
                rows.append((float(parts[1]), float(parts[3]), float(parts[4])))
    rows.sort()
    with open(os.path.join(OUT, "gnear_alpha0.1.txt"), "w") as f:
        f.write("# r G(alpha=0.1,r) dG/dr   (reference tabulated values, 6 significant digits)\n")
        for r, g, d in rows:
            f.write(f"{r!r} {g!r} {d!r}\n")
    return len(rows)


def lj_salt():
    path = os.path.join(REF, "rbmd", "res", "json_files", "LJ_salt", "box_102030", "rbmd.data")
    lo = np.zeros(3)
    hi = np.zeros(3)
    atoms = []
    vel = {}
    section = None
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line:
                continue
            toks = line.split()
            if len(toks) >= 4 and toks[2].endswith("lo") and toks[3].endswith("hi"):
                d = "xyz".index(toks[2][0])
                lo[d], hi[d] = float(toks[0]), float(toks[1])
                continue
            if toks[0] in ("Masses", "Atoms", "Velocities", "Pair", "Bonds"):
                section = toks[0]
                continue
            if section == "Atoms" and len(toks) >= 6:
                im = [int(t) for t in toks[6:9]] if len(toks) >= 9 else [0, 0, 0]
                atoms.append((int(toks[0]), int(toks[1]), float(toks[2]), float(toks[3]),
                              float(toks[4]), float(toks[5]), *im))
            if section == "Velocities" and len(toks) == 4:
                vel[int(toks[0])] = (float(toks[1]), float(toks[2]), float(toks[3]))
    atoms.sort()
    a = np.array(atoms)
    v = np.array([vel.get(int(i), (np.nan,) * 3) for i in a[:, 0]])
    np.savez_compressed(os.path.join(OUT, "lj_salt_6000.npz"), ids=a[:, 0].astype(np.int64),
                        types=a[:, 1].astype(np.int32), q=a[:, 2], x=a[:, 3], y=a[:, 4],
                        z=a[:, 5], img=a[:, 6:9].astype(np.int32), v=v, lo=lo, hi=hi)
    with open(path, "rb") as src, open(os.path.join(OUT, "lj_salt_6000.data.gz"), "wb") as f, \
            gzip.GzipFile(fileobj=f, mode="wb", mtime=0) as dst:
        dst.write(src.read())
    return len(atoms)


def example_json():
    src = os.path.join(REF, "rbmd", "res", "json_files", "LJ_salt", "rbmd.json")
    with open(src, "rb") as f:
        data = f.read()
    with open(os.path.join(OUT, "lj_salt_rbmd.json"), "wb") as f:
        f.write(data)
    return len(data)


if __name__ == "__main__":
    print("gnear rows:", gnear())
    print("lj_salt atoms:", lj_salt())
    print("lj_salt_rbmd.json bytes:", example_json())
    sys.exit(0)
