"""Input and output formats (host code in librbmd_force.so, no GPU).

The LAMMPS data-file reader is checked against the reference's own example input
(tests/golden/lj_salt_6000.data.gz, a copy of rbmd/res/json_files/LJ_salt/box_102030/rbmd.data)
and the arrays the fixture script parsed from it independently
(tests/golden/lj_salt_6000.npz); the atomic and full styles, image flags, Masses and
Pair Coeffs and the error cases on small hand-written files; the dump writer by
parsing its frames back; the JSON loader on an input with the reference's keys,
comments and trailing commas."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _fixture_data(tmp_path):
    p = tmp_path / "rbmd.data"
    with gzip.open(os.path.join(GOLDEN, "lj_salt_6000.data.gz"), "rb") as f:
        p.write_bytes(f.read())
    return str(p)


def test_read_data_reference_example(tmp_path):
    from copy_gpu_amd import formats

    d = formats.read_data(_fixture_data(tmp_path), "charge")
    g = np.load(os.path.join(GOLDEN, "lj_salt_6000.npz"))
    assert d["n"] == 6000 and d["ntypes"] == 2
    assert np.array_equal(g["ids"], np.arange(1, 6001))
    assert np.array_equal(d["lo"], g["lo"]) and np.array_equal(d["hi"], g["hi"])
    assert np.array_equal(d["type"], g["types"] - 1)
    for k in ("q", "x", "y", "z"):
        assert np.array_equal(d[k], g[k]), k
    assert np.array_equal(d["img"].reshape(-1, 3), g["img"])
    assert np.abs(g["img"]).max() == 1  # the example does hold image flags
    assert d["has_velocities"] and d["has_masses"] and d["has_pair_coeffs"]
    assert np.array_equal(np.stack([d["vx"], d["vy"], d["vz"]], 1), g["v"])
    assert np.array_equal(d["mass"], [1.0, 1.0])
    assert np.array_equal(d["eps"], [1.0, 1.0]) and np.array_equal(d["sigma"], [1.0, 1.0])
    # the salt is neutral: 3000 anions of type 1, 3000 cations of type 2
    assert d["q"].sum() == 0 and np.array_equal(np.bincount(d["type"]), [3000, 3000])
    assert np.all(d["q"][d["type"] == 0] == -1) and np.all(d["q"][d["type"] == 1] == 1)


HEAD = """title line
3 atoms
2 atom types

-1.0 4.0 xlo xhi
0 5 ylo yhi
0.5 6.5 zlo zhi

Masses

1 2.5
2 0.75

Pair Coeffs # lj/cut

1 1.2 0.9
2 0.8 1.1

"""


def _write(tmp_path, text):
    p = tmp_path / "t.data"
    p.write_text(text)
    return str(p)


def test_read_data_styles_and_order(tmp_path):
    from copy_gpu_amd import formats

    atomic = HEAD + "Atoms # atomic\n\n3 2 0.1 0.2 0.3\n1 1 1.5 2.5 3.5 0 1 -1\n2 2 3 4 5\n"
    d = formats.read_data(_write(tmp_path, atomic), "atomic")
    assert np.array_equal(d["lo"], [-1, 0, 0.5]) and np.array_equal(d["hi"], [4, 5, 6.5])
    assert np.array_equal(d["type"], [0, 1, 1])
    assert np.array_equal(d["x"], [1.5, 3, 0.1]) and np.array_equal(d["z"], [3.5, 5, 0.3])
    assert np.array_equal(d["img"], [0, 1, -1, 0, 0, 0, 0, 0, 0])
    assert np.array_equal(d["mass"], [2.5, 0.75])
    assert np.array_equal(d["eps"], [1.2, 0.8]) and np.array_equal(d["sigma"], [0.9, 1.1])
    assert not d["has_velocities"]
    full = HEAD + ("Atoms # full\n\n2 7 1 -0.5 1 1 1\n1 7 2 0.5 2 2 2\n3 8 1 0 3 3 3\n\n"
                   "Velocities\n\n3 0.3 0.2 0.1\n1 -1 0 1\n2 5 6 7\n")
    d = formats.read_data(_write(tmp_path, full), "full")
    assert np.array_equal(d["q"], [0.5, -0.5, 0]) and np.array_equal(d["type"], [1, 0, 0])
    assert np.array_equal(d["vx"], [-1, 5, 0.3]) and np.array_equal(d["vz"], [1, 7, 0.1])


@pytest.mark.parametrize("body,style,match", [
    ("Atoms\n\n1 1 0 0 0 0\n1 1 0 1 1 1\n3 1 0 2 2 2\n", "charge", "repeated"),
    ("Atoms\n\n1 1 0 0 0 0\n2 1 0 1 1 1\n", "charge", "lacks"),
    ("Atoms\n\n1 3 0 0 0 0\n2 1 0 1 1 1\n3 1 0 2 2 2\n", "charge", "type out of range"),
    ("Atoms\n\n1 1 0 0 0 0\n2 1 0 1 1 1\n4 1 0 2 2 2\n", "charge", "id out of range"),
    ("Atoms\n\n1 1 0 0 0\n", "charge", "bad atom line"),
    ("Atoms\n\n1 1 0 0 0 0\n", "molecular", "atom_style"),
])
def test_read_data_errors(tmp_path, body, style, match):
    from copy_gpu_amd import formats
    from copy_gpu_amd.force import RBMDError

    with pytest.raises(RBMDError, match=match):
        formats.read_data(_write(tmp_path, HEAD + body), style)


def test_read_data_missing_header(tmp_path):
    from copy_gpu_amd import formats
    from copy_gpu_amd.force import RBMDError

    with pytest.raises(RBMDError, match="header"):
        formats.read_data(_write(tmp_path, "t\n3 atoms\n\nAtoms\n\n1 1 0 0 0 0\n"))
    with pytest.raises(RBMDError, match="cannot open"):
        formats.read_data(str(tmp_path / "absent.data"))


def test_write_dump_frames_parse_back(tmp_path):
    from copy_gpu_amd import formats

    rng = np.random.default_rng(4)
    n = 57
    lo, hi = np.array([0.0, -1.0, 2.0]), np.array([10.0, 19.0, 32.0])
    t = rng.integers(0, 2, n).astype(np.int32)
    frames = [(100 * k, [rng.uniform(lo[d], hi[d], n) for d in range(3)]) for k in (1, 2)]
    p = str(tmp_path / "rbmd.trj")
    for k, (step, xyz) in enumerate(frames):
        formats.write_dump(p, step, lo, hi, t, *xyz, append=k > 0)
    lines = open(p).read().splitlines()
    per = 9 + n
    assert len(lines) == 2 * per
    for k, (step, xyz) in enumerate(frames):
        fr = lines[k * per:(k + 1) * per]
        assert fr[0] == "ITEM: TIMESTEP" and int(fr[1]) == step
        assert fr[2] == "ITEM: NUMBER OF ATOMS" and int(fr[3]) == n
        assert fr[4] == "ITEM: BOX BOUNDS pp pp pp"
        assert np.array_equal(np.array([[float(v) for v in s.split()] for s in fr[5:8]]),
                              np.stack([lo, hi], 1))
        assert fr[8] == "ITEM: ATOMS id type x y z"
        rows = np.array([[float(v) for v in s.split()] for s in fr[9:]])
        assert np.array_equal(rows[:, 0], np.arange(1, n + 1))
        assert np.array_equal(rows[:, 1], t + 1)
        for d in range(3):
            assert np.array_equal(rows[:, 2 + d], xyz[d])  # 17 digits round-trip exactly
    # append=False starts the file again
    formats.write_dump(p, 7, lo, hi, t, *frames[0][1], append=False)
    assert len(open(p).read().splitlines()) == per


def test_load_config_reference_keys(tmp_path):
    from copy_gpu_amd import formats

    data = HEAD + "Atoms # charge\n\n1 1 -1 0 0 1\n2 2 1 1 1 1\n3 2 0 2 2 2\n"
    (tmp_path / "box.data").write_text(data)
    js = """﻿{
  "init_configuration": {
      "read_data": {
      "file": "box.data",
      "unit": "LJ", //LJ ,REAL
      "atom_style": "charge",
      "velocity_type": "GAUSS"
    }
  },
  "hyper_parameters": {
    "force_field": { "type": "LJ/CUT/COUL/LONG" },
    "neighbor": { "type": "VERLET-LIST", "cut_off": 2.5 },
    "coulomb": { "type": "RBE", "coulomb_sample_num": 64, "alpha": 0.3, "kmax": 2 },
    "extend": {
    //  "fix_shake": false
    },
  },
  "execution": {
    "ensemble": "NVT",
    "temperature": [ 1.5, 1.5, 0.25 ],
    "temp_ctrl_type": "BERENDSEN",
    "timestep": 1e-3,
    "num_steps": 40
  },
  "outputs": { "trajectory_out": { "interval": 10 } }
}
"""
    p = tmp_path / "rbmd.json"
    p.write_text(js)
    fc, run, raw = formats.load_config(str(p), coulomb_cut=2.0)
    assert fc.box_lo == (-1.0, 0.0, 0.5) and fc.box_len == (5.0, 5.0, 6.0)
    assert fc.cut_off == 2.5 and fc.coulomb_cut == 2.0 and fc.alpha == 0.3
    assert fc.kspace == "RBE" and fc.coulomb_sample_num == 64 and fc.n_max == 3
    assert fc.eps == (1.2, 0.8) and fc.sigma == (0.9, 1.1)
    assert run["dt"] == 1e-3 and run["num_steps"] == 40
    assert run["temp_ctrl_type"] == "BERENDSEN" and run["temperature"] == (1.5, 1.5, 0.25)
    assert run["trajectory_interval"] == 10
    assert np.array_equal(run["data"]["q"], [-1, 1, 0])
