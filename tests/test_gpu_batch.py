"""GPU tests of batched uneven replicas: a batch step is one launch per phase over
all replicas, and every replica must take bit-identical steps to the same system
advanced alone (same kernels, same k-lists), in graph and eager mode."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261202


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _cfg(L, n, kspace="RBE", P=100, precision=32):
    from copy_gpu_amd import ForceConfig

    # full rows (the default), the pair pass the batched tile pass uses
    return ForceConfig(box_len=(L, L, L), cut_off=2.5, coulomb_cut=5.0, alpha=0.5, skin=0.3,
                       kspace=kspace, coulomb_sample_num=P, precision=precision, n_max=n)


def _systems(sizes, seed0=100):
    from copy_gpu_amd import systems

    out = []
    for r, n in enumerate(sizes):
        x, y, z, q, L = systems.random_ions(n, 0.3, 0.8, seed=seed0 + r)
        vx, vy, vz = systems.velocities(n, 1.0, seed=seed0 + r)
        out.append((x, y, z, q, L, vx, vy, vz))
    return out


def _state(sysr, fdtype):
    x, y, z, q, L, vx, vy, vz = sysr
    n = x.size
    return [_dev(a) for a in (x, y, z, vx, vy, vz)] + \
        [torch.zeros(n, device="cuda", dtype=fdtype) for _ in range(3)] + \
        [torch.zeros(3 * n, device="cuda", dtype=torch.int32)]


def _prime(f, st, P, seed, kspace):
    if kspace == "RBE":
        f.sample_kvectors(P, seed, 0)
    f.Execute(*st[:3], *st[6:9])


def _thermo(f, kind, seed):
    if kind is not None:
        f.set_thermostat(kind, (1.2, 1.2, 0.05), seed=seed, gamma=10.0)


def _run_single(sysr, seed, nsteps, kspace, P, use_graph, proposal=None, thermostat=None):
    from copy_gpu_amd import LJCutCoulKspaceForce

    x, y, z, q, L, vx, vy, vz = sysr
    f = LJCutCoulKspaceForce(_cfg(L, x.size, kspace, P))
    f.Init(_dev(q))
    if proposal is not None:
        proposal.attach(f)
    _thermo(f, thermostat, seed)
    st = _state(sysr, f.fdtype)
    _prime(f, st, P, seed, kspace)
    f.md_run(*st[:9], dt=0.002, nsteps=nsteps, seed=seed, step0=1, img=st[9],
             use_graph=use_graph)
    torch.cuda.synchronize()
    out = [t.cpu().numpy() for t in st]
    if kspace == "RBE":
        m, w, _, _ = f.kspace()
        out += [m, w]
    if thermostat is not None:
        th = f.thermostat_state()
        out += [np.array([th["T"], th["xi"], th["noise_step"]])]
    f.close()
    return out


def _run_batch(syss, seeds, nsteps, kspace, P, use_graph, proposals=None, thermostats=None):
    from copy_gpu_amd.batch import ReplicaBatch

    b = ReplicaBatch([_cfg(s[4], s[0].size, kspace, P) for s in syss])
    states = []
    for r, (sysr, f) in enumerate(zip(syss, b.replicas)):
        f.Init(_dev(sysr[3]))
        if proposals and proposals[r] is not None:
            proposals[r].attach(f)
        if thermostats:
            _thermo(f, thermostats[r], seeds[r])
        st = _state(sysr, f.fdtype)
        _prime(f, st, P, seeds[r], kspace)
        states.append(st)
    b.md_run([tuple(st) for st in states], dt=0.002, nsteps=nsteps, seeds=seeds, step0=1,
             use_graph=use_graph)
    torch.cuda.synchronize()
    outs = []
    for r, (st, f) in enumerate(zip(states, b.replicas)):
        o = [t.cpu().numpy() for t in st]
        if kspace == "RBE":
            m, w, _, _ = f.kspace()
            o += [m, w]
        if thermostats and thermostats[r] is not None:
            th = f.thermostat_state()
            o += [np.array([th["T"], th["xi"], th["noise_step"]])]
        outs.append(o)
    stats = [f.stats() for f in b.replicas]
    b.close()
    return outs, stats


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_rbe_equals_single_runs(gpu, use_graph):
    sizes = [2048, 5000, 3100, 9000, 2600]
    _log(f"case: batch of {len(sizes)} uneven RBE replicas {sizes}, 40 steps, graph={use_graph}")
    syss = _systems(sizes)
    seeds = [SEED + 7 * r for r in range(len(sizes))]
    outs, stats = _run_batch(syss, seeds, 40, "RBE", 100, use_graph)
    assert all(s["builds"] >= 2 for s in stats)  # the step includes real rebuilds
    for r, sysr in enumerate(syss):
        ref = _run_single(sysr, seeds[r], 40, "RBE", 100, use_graph=True)
        for a, b_ in zip(outs[r], ref):
            assert np.array_equal(a, b_), f"replica {r} differs"


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_thermostats_equal_single_runs(gpu, use_graph):
    """Replicas with different NVT thermostats (and one without) in one batch: each
    takes bit-identical steps to the same handle advanced alone by md_run with its
    thermostat (ExecutionNVT::Solve + UpdateVelocityByTempConType,
    rbmd/framework/run/ExecutionNVT.cpp:133-159, 256-291), including the final
    temperature, Nose-Hoover friction and Langevin noise step."""
    sizes = [2048, 3500, 2300, 4100, 2700]
    kinds = ["BERENDSEN", None, "NOSE_HOOVER", "LANGEVIN", "TEMP_RESCALE"]
    _log(f"case: batch thermostats {kinds} on replicas {sizes}, 20 steps, graph={use_graph}")
    syss = _systems(sizes, seed0=500)
    seeds = [SEED + 11 * r for r in range(len(sizes))]
    outs, _ = _run_batch(syss, seeds, 20, "RBE", 100, use_graph, thermostats=kinds)
    for r, sysr in enumerate(syss):
        ref = _run_single(sysr, seeds[r], 20, "RBE", 100, use_graph=True, thermostat=kinds[r])
        assert len(outs[r]) == len(ref)
        for a, b_ in zip(outs[r], ref):
            assert np.array_equal(a, b_), f"replica {r} ({kinds[r]}) differs"
    # the thermostats acted: velocities of the Berendsen replica differ from NVE
    nve = _run_single(syss[0], seeds[0], 20, "RBE", 100, use_graph=True)
    assert not np.array_equal(outs[0][3], nve[3])


def test_batch_lj_only_equals_single_runs(gpu):
    sizes = [2500, 4100, 2048]
    _log(f"case: batch of LJ + real-space replicas {sizes}, 30 steps")
    syss = _systems(sizes, seed0=300)
    seeds = [0] * len(sizes)
    outs, _ = _run_batch(syss, seeds, 30, "NONE", 0, True)
    for r, sysr in enumerate(syss):
        ref = _run_single(sysr, 0, 30, "NONE", 0, use_graph=False)
        for a, b_ in zip(outs[r], ref):
            assert np.array_equal(a, b_), f"replica {r} differs"


def test_batch_with_learned_proposals(gpu):
    from copy_gpu_amd import proposal as PR

    sizes = [3000, 2048, 6000]
    _log(f"case: batch with per-replica learned proposals {sizes}, 25 steps")
    syss = _systems(sizes, seed0=500)
    seeds = [11, 12, 13]
    rng = np.random.default_rng(0)
    props = []
    for r, s in enumerate(syss):
        if r == 1:
            props.append(None)  # this replica keeps the analytic sampler
            continue
        L = s[4]
        Mg = 8
        m = PR.lattice_grid(Mg)
        d = PR.analytic_weight(m, (L, L, L), 0.5) * rng.uniform(0.5, 2.0, len(m))
        props.append(PR.Proposal(Mg, d, 0.9, 0.5, (L, L, L)))
    outs, _ = _run_batch(syss, seeds, 25, "RBE", 100, True, props)
    for r, sysr in enumerate(syss):
        ref = _run_single(sysr, seeds[r], 25, "RBE", 100, use_graph=True, proposal=props[r])
        for a, b_ in zip(outs[r], ref):
            assert np.array_equal(a, b_), f"replica {r} differs"


def test_batch_autotune_then_run(gpu):
    """The batch autotuner times one shape policy for all replicas as batched launches;
    it leaves velocities unchanged, and the tuned batch still advances every replica
    with forces equal to a fresh evaluation at the final positions."""
    from copy_gpu_amd.batch import ReplicaBatch

    sizes = [2048, 7000, 3500, 12000]
    _log(f"case: batch autotune on replicas {sizes}, then 20 steps")
    syss = _systems(sizes, seed0=700)
    b = ReplicaBatch([_cfg(s[4], s[0].size) for s in syss])
    states = []
    for r, (sysr, f) in enumerate(zip(syss, b.replicas)):
        f.Init(_dev(sysr[3]))
        st = _state(sysr, f.fdtype)
        _prime(f, st, 100, 40 + r, "RBE")
        states.append(st)
    v0 = [[t.clone() for t in st[3:6]] for st in states]
    shapes = b.autotune([tuple(st) for st in states], reps=2)
    _log(f"  tuned shapes {shapes}")
    assert shapes["tuned_cost_us"] > 0
    for st, v in zip(states, v0):
        for a, c in zip(st[3:6], v):
            assert torch.equal(a, c)
    b.md_run([tuple(st) for st in states], dt=0.002, nsteps=20, seeds=[40 + r for r in
             range(len(sizes))], step0=1)
    torch.cuda.synchronize()
    from copy_gpu_amd import systems

    for r, (st, f, sysr) in enumerate(zip(states, b.replicas, syss)):
        F = torch.stack(st[6:9], 1).double()
        assert torch.isfinite(F).all()
        # the last step (20) used the draws of step 20; the handle now holds step 21's
        L = sysr[4]
        m, _ = systems.sample_kvectors_host(0.5, np.array([L] * 3), 100, 40 + r, 20)
        f.set_kvectors(_dev(m.reshape(-1), torch.int32))
        Fc = [torch.zeros_like(st[6]) for _ in range(3)]
        f.Execute(*st[:3], *Fc)
        Fc = torch.stack(Fc, 1).double()
        assert float(torch.linalg.norm(F - Fc) / torch.linalg.norm(Fc)) < 1e-5
    b.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_step_longer_than_half_skin_names_the_replica(gpu, use_graph):
    """A batched run ends with one gathered copy of every replica's statistics and
    flags; the long-step check reaches the caller through it: one atom of replica 1
    moving 0.2 in a 0.002 step (half the 0.3 skin is 0.15) makes the run report
    replica 1, and a batch of the same replicas without it runs clean."""
    from copy_gpu_amd import RBMDError

    _log(f"case: batch long-step flag, 2 replicas, graph={use_graph}")
    syss = _systems([3000, 4000], seed0=300)
    outs, _ = _run_batch(syss, [5, 6], 6, "RBE", 100, use_graph)
    assert all(np.isfinite(o[0]).all() for o in outs)
    x, y, z, q, L, vx, vy, vz = syss[1]
    vx = vx.copy()
    vx[7] = 100.0
    syss[1] = (x, y, z, q, L, vx, vy, vz)
    with pytest.raises(RBMDError, match=r"replica 1\)"):
        _run_batch(syss, [5, 6], 6, "RBE", 100, use_graph)
    torch.cuda.synchronize()
