"""GPU run of the reference's own example input: the LJ_salt data file
(tests/golden/lj_salt_6000.data.gz, 6000 +/-1 ions, box 10 x 20 x 30, two types) read
by the library's read_data, FP64 forces with full Ewald k-space against the oracle,
then a dump frame of the configuration written and parsed back."""
import gzip
import os
import sys

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_reference_example_input_fp64_ewald(orc, gpu, tmp_path):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce, formats
    from test_gpu_md import _halfspace_lattice

    _log("case: LJ_salt example data file, FP64 Ewald forces, 6000 ions")
    path = tmp_path / "rbmd.data"
    with gzip.open(os.path.join(GOLDEN, "lj_salt_6000.data.gz"), "rb") as f:
        path.write_bytes(f.read())
    d = formats.read_data(str(path), "charge")
    lo, Lb = d["lo"], d["hi"] - d["lo"]
    alpha, tol, rc = 0.5476, 1e-12, 2.5  # "alpha" of the example's JSON
    f = LJCutCoulKspaceForce(ForceConfig(
        box_len=tuple(Lb), box_lo=tuple(lo), cut_off=rc, coulomb_cut=rc, skin=0.3,
        alpha=alpha, kspace="EWALD", precision=64, eps=tuple(d["eps"]),
        sigma=tuple(d["sigma"]), n_max=d["n"], ewald_tol=tol))
    dev = lambda a, t=torch.float64: torch.as_tensor(a).to("cuda", t)  # noqa: E731
    f.Init(dev(d["q"]), dev(d["type"], torch.int32))
    X, Y, Z = dev(d["x"]), dev(d["y"]), dev(d["z"])
    F = [torch.zeros(d["n"], device="cuda", dtype=torch.float64) for _ in range(3)]
    f.Execute(X, Y, Z, *F)
    torch.cuda.synchronize()
    Fg = np.stack([t.cpu().numpy() for t in F], 1)
    m, w = _halfspace_lattice(Lb, alpha, tol)
    assert f.stats()["ewald_k"] == m.shape[0]
    Fp, _ = orc.pair_forces(d["x"], d["y"], d["z"], d["q"], lo, Lb, rc, rc, alpha,
                            types=d["type"], eps=tuple(d["eps"]), sigma=tuple(d["sigma"]))
    Fk, _ = orc.kspace_forces(d["x"] - lo[0], d["y"] - lo[1], d["z"] - lo[2], d["q"], Lb,
                              alpha, m, w)
    Fo = Fp + Fk
    rel = np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo)
    _log(f"  {m.shape[0]} half-space k-vectors; relative force error {rel:.2e}")
    assert rel < 1e-12
    # the configuration out as a dump frame, parsed back
    trj = str(tmp_path / "rbmd.trj")
    formats.write_dump(trj, 0, d["lo"], d["hi"], d["type"], d["x"], d["y"], d["z"],
                       append=False)
    rows = np.loadtxt(trj, skiprows=9)
    assert np.array_equal(rows[:, 2], d["x"]) and np.array_equal(rows[:, 4], d["z"])
