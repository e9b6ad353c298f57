"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared symbol, Philox known answers, deterministic generators."""
import os
import re
import subprocess

import numpy as np

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rbmd_force.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbmd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from copy_gpu_amd import _lib

    lib = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rbmd_\w+)", out))
    assert set(names) <= exported, set(names) - exported
    # the ctypes table binds exactly the declared entry points
    assert sorted(_lib.exported_symbols()) == names


def test_library_has_gfx950_code_object():
    from copy_gpu_amd import _lib

    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors (Random123 kat_vectors)."""
    from copy_gpu_amd import systems

    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = systems.philox4x32(ctr, key)
        assert tuple(int(v) for v in got) == want, [hex(int(v)) for v in got]


def test_random_ions_generator(orc):
    from copy_gpu_amd import systems

    x, y, z, q, Lb = systems.random_ions(3000, 0.3, 0.8, seed=5)
    assert abs(Lb - (3000 / 0.3) ** (1 / 3)) < 1e-12
    for a in (x, y, z):
        assert a.min() >= 0 and a.max() < Lb
    assert q.sum() == 0 and set(np.unique(q)) == {-1.0, 1.0}
    Lv = np.array([Lb] * 3)
    offs, _ = orc.neighbors_brute(x, y, z, Lv, 0.8 - 1e-9)
    assert offs[-1] == 0  # minimum separation respected
    x2, y2, z2, q2, _ = systems.random_ions(3000, 0.3, 0.8, seed=5)
    assert np.array_equal(x, x2) and np.array_equal(q, q2)
    x3, *_ = systems.random_ions(3000, 0.3, 0.8, seed=6)
    assert not np.array_equal(x, x3)


def test_velocities_generator():
    from copy_gpu_amd import systems

    n = 5000
    vx, vy, vz = systems.velocities(n, 1.0, seed=3)
    assert abs(vx.sum()) < 1e-9 and abs(vy.sum()) < 1e-9 and abs(vz.sum()) < 1e-9
    T = (vx @ vx + vy @ vy + vz @ vz) / (3 * n - 3)
    assert abs(T - 1.0) < 1e-12
    # Gaussian marginals
    assert abs(np.mean(vx ** 4) / np.mean(vx ** 2) ** 2 - 3.0) < 0.25


def test_fcc_generator():
    from copy_gpu_amd import systems

    x, y, z, Lb = systems.fcc(10, 0.8442)
    assert x.size == 4000
    assert abs(Lb - 10 * (4 / 0.8442) ** (1 / 3)) < 1e-12
    assert abs(4000 / Lb ** 3 - 0.8442) < 1e-12


def test_langevin_noise_stream():
    """The Langevin noise vector: Box-Muller on Philox counters (0x4C414E47, j, step_lo,
    step_hi), checked against the same transform of the Philox words in numpy, and its
    moments over many steps (zero mean, unit variance, uncorrelated components)."""
    from copy_gpu_amd import systems

    def u53(a, b):
        return ((int(a) >> 5) * 67108864.0 + (int(b) >> 6)) / 9007199254740992.0

    seed = 0x1234_5678_9ABC_DEF0
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for step in (0, 1, 7, (5 << 32) + 3):
        r0 = systems.philox4x32((0x4C414E47, 0, step & 0xFFFFFFFF, step >> 32), key)
        r1 = systems.philox4x32((0x4C414E47, 1, step & 0xFFFFFFFF, step >> 32), key)
        a = np.sqrt(-2.0 * np.log(1.0 - u53(r0[0], r0[1])))
        t = 2 * np.pi * u53(r0[2], r0[3])
        b = np.sqrt(-2.0 * np.log(1.0 - u53(r1[0], r1[1])))
        t2 = 2 * np.pi * u53(r1[2], r1[3])
        want = np.array([a * np.cos(t), a * np.sin(t), b * np.cos(t2)])
        got = systems.langevin_noise(seed, step)
        assert np.allclose(got, want, rtol=1e-14, atol=1e-15), (got, want)
    G = np.array([systems.langevin_noise(11, s) for s in range(20000)])
    assert np.abs(G.mean(0)).max() < 4.5 / np.sqrt(20000)
    assert np.abs(G.var(0) - 1).max() < 0.05
    C = np.corrcoef(G.T)
    assert np.abs(C - np.eye(3)).max() < 0.04
    # the stream depends on the seed
    assert not np.array_equal(systems.langevin_noise(11, 3), systems.langevin_noise(12, 3))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 ends with a message and no JSON line."""
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
