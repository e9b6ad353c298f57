#!/usr/bin/env python3
"""Benchmark: atom-steps/s of the charged-LJ + Random Batch Ewald MD step on one MI355X,
at fixed force RMS error against FP64 Ewald, learned vs plain RBE.

Workload (default): 100,000 +/-1 ions, uniform random with minimum separation
0.8, density 0.3 (L = 69.34), LJ cutoff 2.5, erfc real-space cutoff 5.0,
alpha 0.5, Verlet skin 0.3, T = 1, dt = 0.002; FP32 pair and k-space maths, FP64
positions, velocities, integrator.  The k-space proposal network (4-64-64-1 MLP)
is trained on the GPU on FP64 |rho(k)|^2 of 50 MD snapshots, then learned P = 50
is compared with the analytic sampler at P = 50/100/200/400: force RMS error
against FP64 full Ewald (5 held-out configurations x 20 k-lists each) and
atom-steps/s of each.  The headline value is the learned P = 50 run.  A step =
kick, drift, list check (and rebuild when needed), k-sampling, structure factor,
pair pass, k-space force, kick -- captured in HIP graphs (32 steps and the
run's remainder) before the timed region and replayed.

Prints ONE JSON line on stdout (logs go to stderr).  Also reported:
  * roofline of the pair pass (the dominant kernel): algorithmic bytes
    4 E + 40 N per launch over its mean launch time (20 launches back to back
    between one pair of HIP events on its stream; E = full-list entries at
    r_list), against 8 TB/s;
  * the CPU baseline: the reference-faithful step on the host cores for a bounded
    sample of steps of the same workload;
  * beside the headline: the 32,768-ion plain-RBE workload (timed steps, pair-pass
    roofline, Berendsen NVT) and the pair-pass roofline of the 1,048,576-ion one.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s at fixed force RMS error vs FP64 Ewald; learned vs plain RBE"
HBM_PEAK_GBS = 8000.0
SEED = 20261202

WORKLOADS = {
    "c1": dict(n=4000, fcc=10, rho=0.8442, rc_lj=2.5, rc_coul=0.0, alpha=0.0, P=0, dt=0.005,
               steps=1000, kspace="NONE", precision=64, ewald=False,
               desc="plain LJ fluid, FP64: 4000 atoms on an FCC lattice at density 0.8442, "
                    "rc 2.5, skin 0.3, T 1.0, dt 0.005, no charges"),
    "c2": dict(n=32768, rho=0.3, rc_lj=2.5, rc_coul=5.0, alpha=0.5, P=100, dt=0.002,
               steps=500, desc="charged LJ electrolyte, plain RBE: 32768 ions, density 0.3, "
                           "rc_LJ 2.5, rc_coul 5.0, alpha 0.5, P=100, dt 0.002"),
    # no FP64 full-Ewald reference at this size: alpha 0.3 to 1e-14 needs about 7e6
    # half-space k-vectors for 1,048,576 ions (about 7e12 FP64 phases per configuration)
    "c3": dict(n=1048576, rho=0.3, rc_lj=2.5, rc_coul=6.0, alpha=0.3, P=500, dt=0.002,
               steps=200, ewald=False, desc="large charged LJ, plain RBE: 1048576 ions, density 0.3, "
                           "rc_LJ 2.5, rc_coul 6.0, alpha 0.3, P=500, dt 0.002"),
    "c4": dict(n=100000, rho=0.3, rc_lj=2.5, rc_coul=5.0, alpha=0.5, P=50, dt=0.002,
               steps=200, learned=True, Mg=32, equilibrate=500, snapshots=50, snap_every=10,
               plain_P=(50, 100, 200, 400),
               desc="learned k-space proposal: 100000 ions, density 0.3, rc_LJ 2.5, "
                    "rc_coul 5.0, alpha 0.5; MLP 4-64-64-1 trained on FP64 |rho(k)|^2 of 50 "
                    "MD snapshots on |m_d| <= 32; learned P=50 vs analytic P=50/100/200/400"),
    "c6": dict(n=10_000_000, rho=0.3, rc_lj=2.5, rc_coul=6.0, alpha=0.3, P=500, dt=0.002,
               steps=40, ewald=False, cpu=False,
               desc="capacity: the large charged-LJ system scaled to 10,000,000 ions, density "
                    "0.3, rc_LJ 2.5, rc_coul 6.0, alpha 0.3, P=500, dt 0.002"),
    "c5": dict(n=0, replicas=64, n_min=2048, n_max=65536, rho=0.3, rc_lj=2.5, rc_coul=5.0,
               alpha=0.5, P=100, dt=0.002, steps=300, batch=True,
               desc="64 uneven replicas, N_r = floor(2048 * 32^u_r) seeded (log-uniform on "
                    "[2048, 65536]), density 0.3, rc_LJ 2.5, rc_coul 5.0, alpha 0.5, P=100 "
                    "learned-proposal k-samples per replica, dt 0.002; one batched launch "
                    "per phase, HIP graph per step, vs the same replicas run one after another"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name, key="bytes_per_launch"):
    """A per-launch figure of a kernel from a committed rocprofv3 --pmc summary
    (profiles/<name>_traffic.txt, line '<key>: <value>'), else None."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


ENGINE_CLOCK_HZ = 2.4e9  # MI355X peak engine clock; one VALU wave-instruction per CU-cycle


def cpu_baseline(wl, x, y, z, q, vx, vy, vz, L, target_s):
    import oracle

    n, P = wl["n"], wl["P"]
    threads = max(1, min(16, os.cpu_count() or 1))
    real = np.float64 if wl.get("precision", 32) == 64 else np.float32
    pos = np.ascontiguousarray(np.stack([x, y, z], 1), real)
    vel = np.ascontiguousarray(np.stack([vx, vy, vz], 1), real)
    frc = np.zeros_like(pos)
    if wl.get("kspace", "RBE") == "NONE":
        q = None
    from copy_gpu_amd import systems

    def klist(steps, s0):
        if P == 0:
            return None
        return np.concatenate([systems.sample_kvectors_host(wl["alpha"], [L] * 3, P, SEED,
                                                            s0 + s)[0] for s in range(steps)])

    lo, Lb = np.zeros(3), np.array([L] * 3)
    t = time.perf_counter()
    oracle.faithful_run(pos, vel, frc, q, lo, Lb, wl["rc_lj"], wl["rc_coul"], wl["alpha"], P,
                        klist(1, 0), 1, wl["dt"], threads)
    t1 = time.perf_counter() - t
    steps = int(max(1, min(2000, target_s / max(t1, 1e-6))))
    m = klist(steps, 1)
    t = time.perf_counter()
    rc = oracle.faithful_run(pos, vel, frc, q, lo, Lb, wl["rc_lj"], wl["rc_coul"], wl["alpha"],
                             P, m, steps, wl["dt"], threads)
    el = time.perf_counter() - t
    if rc != 0:
        log("cpu baseline: list slots overflowed (reference fixed-stride list)")
    # the same path on one thread (the reference's serial VTK-m device), a shorter sample
    steps1 = int(max(1, min(2000, target_s / 3.0 / max(t1 * threads, 1e-6))))
    m1 = klist(steps1, 1 + steps)
    t = time.perf_counter()
    oracle.faithful_run(pos, vel, frc, q, lo, Lb, wl["rc_lj"], wl["rc_coul"], wl["alpha"], P,
                        m1, steps1, wl["dt"], 1)
    el1 = time.perf_counter() - t
    prec = "FP64" if real == np.float64 else "FP32"
    what = ("N*P structure-factor arrays, analytic sampler at the same P" if P else
            "LJ only")
    return {"value": n * steps / el, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "single_thread_value": n * steps1 / el1, "host_cpus": os.cpu_count(),
            "sample": f"{steps} reference-faithful {prec} steps (per-step counting sort by "
                      f"cell, list at rc without skin, {what}) of the same {n}-atom "
                      f"workload, OpenMP over atoms, {el:.1f} s; then {steps1} steps on one "
                      f"thread, {el1:.1f} s"}


# --half-shell: FP32 handles use the opt-in half-shell pair pass (each pair once)
HALF_SHELL = False


def make_force(wl, L, P, kspace="RBE", precision=32, tol=1e-14):
    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce

    return LJCutCoulKspaceForce(ForceConfig(
        box_len=(L, L, L), cut_off=wl["rc_lj"], coulomb_cut=wl["rc_coul"], skin=0.3,
        alpha=wl["alpha"], kspace=kspace, coulomb_sample_num=max(P, 1), precision=precision,
        n_max=wl["n"], ewald_tol=tol, half_shell=HALF_SHELL))


class State:
    """Device MD state (FP64 positions and velocities, FP32 forces)."""

    def __init__(self, X, Y, Z, VX, VY, VZ, F):
        self.X, self.Y, self.Z, self.VX, self.VY, self.VZ, self.F = X, Y, Z, VX, VY, VZ, F

    def clone(self):
        return State(*(t.clone() for t in (self.X, self.Y, self.Z, self.VX, self.VY, self.VZ)),
                     [t.clone() for t in self.F])

    def run(self, f, wl, nsteps, step0, use_graph=True):
        f.md_run(self.X, self.Y, self.Z, self.VX, self.VY, self.VZ, *self.F, dt=wl["dt"],
                 nsteps=nsteps, seed=SEED, step0=step0, use_graph=use_graph)


# Steps per cached HIP graph in the library: a K-step run replays the GRAPH_STEPS-step
# graph K // GRAPH_STEPS times and the graph of the K % GRAPH_STEPS remainder once,
# each captured on first use.
GRAPH_STEPS = 32


def graph_warmup(W, K):
    """At least W warm-up steps whose run captures every graph a K-step run replays:
    the GRAPH_STEPS-step graph (when K >= GRAPH_STEPS) and the remainder's graph,
    so no capture happens inside the timed region."""
    r = K % GRAPH_STEPS
    need = (GRAPH_STEPS if K >= GRAPH_STEPS else 0) + r
    extra = max(0, W - need)
    return need + GRAPH_STEPS * ((extra + GRAPH_STEPS - 1) // GRAPH_STEPS)


def timed(f, st, wl, W, K, step, use_graph):
    """Warm-up steps (with graphs, graph_warmup(W, K) of them, so every graph the
    timed run replays is captured and instantiated before the clock starts), then K
    timed MD steps; returns (seconds, list builds, next step)."""
    import torch

    Wr = graph_warmup(W, K) if use_graph else W
    st.run(f, wl, Wr, step, use_graph)
    step += Wr
    torch.cuda.synchronize()
    b0 = f.stats()["builds"]
    t0 = time.perf_counter()
    st.run(f, wl, K, step, use_graph)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return el, f.stats()["builds"] - b0, step + K


def nvt_leg(f, st, wl, W, K, step, use_graph):
    """The same MD steps from the same configuration under the reference's default NVT
    thermostat: Berendsen, "temperature": [1.0, 1.0, 0.5] (res/json_files/LJ_salt/rbmd.json:39-41)."""
    import torch

    s2 = st.clone()
    f.set_thermostat("BERENDSEN", (1.0, 1.0, 0.5))
    el, rebuilds, step = timed(f, s2, wl, W, K, step, use_graph)
    T = f.thermostat_state()["T"]
    f.set_thermostat("NONE")
    n = s2.X.numel()
    del s2
    torch.cuda.synchronize()
    return {"thermostat": "BERENDSEN", "t_target": 1.0, "t_damp": 0.5,
            "value": n * K / el, "unit": "atom-steps/s", "ms_per_step": el * 1e3 / K,
            "steps": K, "list_builds": rebuilds, "temperature_last_step": T,
            "note": "four kernels per step instead of three: opening kick (summing the "
                    "step before's temperature and applying the rescale), structure factor, "
                    "tile pass, closing kick (per-workgroup sums of m v^2)"}


def kernel_times(f, st, wl, Kt, step):
    """Per-kernel device ms per step from HIP events over Kt eager steps, and the
    pair-pass roofline numbers."""
    import torch

    f.set_timing(True)
    st.run(f, wl, Kt, step, use_graph=False)
    torch.cuda.synchronize()
    names = ["check", "bin", "scan", "scatter", "cellsort", "gather", "sample", "rho",
             "rho_final", "pair", "kforce"]
    per = {}
    for nm in names:
        ms, cnt = f.timing(nm)
        per[nm] = {"mean_ms": ms, "launches": cnt}
    # the pair pass alone at the final positions: 20 launches back to back between
    # one pair of HIP events on the launch stream (the library's time_pair), the
    # time a kernel trace of the same launches gives
    f.set_timing(False)
    pair_cnt = 20
    pair_ms = f.time_pair(st.X, st.Y, st.Z, *st.F, reps=pair_cnt)
    torch.cuda.synchronize()
    E = f.stats()["entries"]
    n = wl["n"]
    pair_bytes = 4.0 * E + 40.0 * n
    per_step_ms = {k: v["mean_ms"] * v["launches"] / Kt for k, v in per.items()}
    log(f"pair pass without rebuild: {pair_ms * 1e3:.2f} us over {pair_cnt} launches")
    log("per-step device ms: " + ", ".join(f"{k} {v:.4f}" for k, v in per_step_ms.items()))
    return per_step_ms, pair_bytes, pair_ms, E, step + Kt


def pair_pass_at_scale(wname, reps=20):
    """Pair-pass roofline of another workload at fixed positions (its starting
    configuration): list built once, then reps pair-pass launches back to back
    between one pair of HIP events."""
    import torch

    from copy_gpu_amd import systems

    wl = WORKLOADS[wname]
    n = wl["n"]
    x, y, z, q, L = systems.random_ions(n, wl["rho"], 0.8, seed=SEED)
    t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to("cuda:0", torch.float64)  # noqa
    X, Y, Z, Q = t64(x), t64(y), t64(z), t64(q)
    F = [torch.zeros(n, device="cuda:0", dtype=torch.float32) for _ in range(3)]
    f = make_force(wl, L, wl["P"])
    f.Init(Q)
    f.sample_kvectors(wl["P"], SEED, 0)
    pair_ms = f.time_pair(X, Y, Z, *F, reps=reps)
    cnt = reps
    torch.cuda.synchronize()
    E = f.stats()["entries"]
    f.close()
    nbytes = 4.0 * E + 40.0 * n
    achieved = nbytes / (pair_ms * 1e-3) / 1e9
    valu = read_traffic(f"pair_{wname}", "valu_wave_instructions_per_launch")
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    out = {"workload": wl["desc"], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(f"pair_{wname}"),
           "algorithmic_bytes_per_launch": nbytes, "mean_launch_ms": pair_ms,
           "launches": cnt, "list_entries": E,
           "valu_issue_frac": valu / (pair_ms * 1e-3 * n_cu * ENGINE_CLOCK_HZ) if valu else None}
    log(f"pair pass at {wname}: {pair_ms * 1e3:.1f} us, {achieved:.0f} GB/s "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of HBM peak)")
    return out


def ewald_refs(wl, L, Q, snaps):
    """FP64 full-Ewald total forces and the Ewald k-space part for each configuration."""
    import torch

    n = wl["n"]
    ew = make_force(wl, L, 0, kspace="EWALD", precision=64, tol=1e-14)
    ew.Init(Q)
    nk = make_force(wl, L, 0, kspace="NONE", precision=64)
    nk.Init(Q)
    refs = []
    for (X, Y, Z) in snaps:
        Fe = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
        Fr = [torch.zeros(n, device="cuda", dtype=torch.float64) for _ in range(3)]
        ew.Execute(X, Y, Z, *Fe)
        nk.Execute(X, Y, Z, *Fr)
        Fe_t = torch.stack(Fe, 1)
        refs.append((Fe_t, Fe_t - torch.stack(Fr, 1)))
    nkv = ew.stats()["ewald_k"]
    ew.close()
    nk.close()
    return refs, nkv


def force_error(f, P, snaps, refs, lists):
    """eps_F over configurations x k-lists: RMS of F - F_Ewald relative to the
    total Ewald force and to its k-space part (same real-space part, so the
    difference is the k-space estimator's error)."""
    import torch

    num = den = denk = 0.0
    n = snaps[0][0].numel()
    F = [torch.zeros(n, device="cuda", dtype=torch.float32) for _ in range(3)]
    for c, ((X, Y, Z), (Fe, Fek)) in enumerate(zip(snaps, refs)):
        for s in range(lists):
            f.sample_kvectors(P, SEED + 1, 1000 * c + s)
            f.Execute(X, Y, Z, *F)
            num += float(((torch.stack(F, 1).double() - Fe) ** 2).sum())
            den += float((Fe ** 2).sum())
            denk += float((Fek ** 2).sum())
    return (num / den) ** 0.5, (num / denk) ** 0.5


def learned_eval(wl, L, Q, st, step, args):
    """Train the proposal on FP64 spectra of MD snapshots, then measure eps_F and
    atom-steps/s of learned P vs analytic P values on held-out configurations."""
    import torch

    from copy_gpu_amd import proposal as PR

    n, alpha, Mg = wl["n"], wl["alpha"], wl["Mg"]
    f = make_force(wl, L, 100)
    f.Init(Q)
    f.sample_kvectors(100, SEED, 0)
    f.Execute(st.X, st.Y, st.Z, *st.F)
    st.run(f, wl, wl["equilibrate"], step)
    step += wl["equilibrate"]
    snaps = []
    log(f"trajectory: {wl['equilibrate']} steps, then {wl['snapshots']} snapshots every "
        f"{wl['snap_every']} steps (analytic RBE, P=100)")
    for i in range(wl["snapshots"]):
        st.run(f, wl, wl["snap_every"], step)
        step += wl["snap_every"]
        snaps.append((st.X.clone(), st.Y.clone(), st.Z.clone()))
    torch.cuda.synchronize()
    t = time.perf_counter()
    S2 = PR.spectrum(snaps, Q, (L, L, L), alpha, Mg)
    torch.cuda.synchronize()
    t_spec = time.perf_counter() - t
    log(f"FP64 spectra of {len(snaps)} snapshots on {(2 * Mg + 1) ** 3} lattice points: "
        f"{t_spec:.2f} s")
    del snaps
    t = time.perf_counter()
    prop, mlp, hist = PR.learned_proposal(S2, (L, L, L), alpha, Mg, lam=0.9)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    log(f"training: loss {hist[0]:.4e} -> {hist[-1]:.4e} in {t_train:.2f} s")
    m = PR.lattice_grid(Mg)
    k = PR.kvecs(m, (L, L, L))
    k2 = (k * k).sum(1)
    a = PR.analytic_weight(m, (L, L, L), alpha)
    opt = np.where(k2 > 0, a * np.sqrt(S2 / np.where(k2 > 0, k2, 1.0)), 0.0)
    va = PR.variance_factor(S2, None, (L, L, L), alpha, Mg)
    g_learned = va / PR.variance_factor(S2, prop.density, (L, L, L), alpha, Mg, prop.lam)
    g_opt = va / PR.variance_factor(S2, opt, (L, L, L), alpha, Mg, 1.0)
    log(f"expected variance gain on the training spectra: learned {g_learned:.4f}, "
        f"optimal k-only proposal {g_opt:.4f}")

    # held-out configurations: 5 more snapshots, 20 steps apart
    snaps = []
    for i in range(5):
        st.run(f, wl, 20, step)
        step += 20
        snaps.append((st.X.clone(), st.Y.clone(), st.Z.clone()))
    f.close()
    refs, nkv = ([], 0) if args.skip_ewald else ewald_refs(wl, L, Q, snaps)
    if refs:
        log(f"FP64 Ewald references: {len(snaps)} configurations, {nkv} half-space k-vectors")

    runs = []
    K = args.steps if args.steps is not None else wl["steps"]
    settings = [("learned", wl["P"])] + [("analytic", p) for p in wl["plain_P"]]
    base = st.clone()
    for kind, P in settings:
        g = make_force(wl, L, P)
        g.Init(Q)
        if kind == "learned":
            prop.attach(g)
        eps = epsk = None
        if refs:
            eps, epsk = force_error(g, P, snaps, refs, 20)
        s2 = base.clone()
        g.sample_kvectors(P, SEED, step)
        g.Execute(s2.X, s2.Y, s2.Z, *s2.F)
        el, rebuilds, _ = timed(g, s2, wl, args.warmup, K, step + 1, not args.no_graph)
        rate = n * K / el
        log(f"{kind:8s} P={P:4d}: eps_F {eps if eps is None else f'{eps:.4e}'} "
            f"(k-space part {epsk if epsk is None else f'{epsk:.4e}'}), "
            f"{rate:.4e} atom-steps/s, {el * 1e3 / K:.4f} ms/step")
        runs.append({"sampler": kind, "P": P, "eps_F": eps, "eps_F_kspace_part": epsk,
                     "atom_steps_per_s": rate, "ms_per_step": el * 1e3 / K,
                     "list_builds": rebuilds})
        if kind == "learned":
            learned_state = (g, s2)
        else:
            g.close()
    out = {"runs": runs, "train": {
        "equilibration_steps": wl["equilibrate"],
        "snapshots": wl["snapshots"], "snapshot_spacing_steps": wl["snap_every"],
        "grid": f"|m_d| <= {Mg} ({(2 * Mg + 1) ** 3} lattice points)",
        "samples": int((Mg + 1) ** 3 - 1), "adam_iters": 2400,
        "loss_initial": hist[0], "loss_final": hist[-1],
        "spectra_s": t_spec, "train_s": t_train, "mixture_lambda": prop.lam},
        "expected_variance_gain": g_learned, "optimal_konly_variance_gain": g_opt}
    # analytic P giving the learned eps_F (eps^2 ~ c / P, least squares over the
    # analytic runs), its throughput by log-log interpolation, and the ratio
    an = [r for r in runs if r["sampler"] == "analytic"]
    lr = runs[0]
    if lr["eps_F"] is not None:
        Ps = np.array([r["P"] for r in an], float)
        e2 = np.array([r["eps_F"] ** 2 for r in an])
        c = float((e2 * Ps).mean())
        P_eq = c / lr["eps_F"] ** 2
        rates = np.array([r["atom_steps_per_s"] for r in an])
        rate_eq = float(np.exp(np.interp(np.log(P_eq), np.log(Ps), np.log(rates))))
        out.update({"analytic_P_at_learned_eps": P_eq,
                    "analytic_atom_steps_per_s_at_learned_eps": rate_eq,
                    "speedup_at_equal_eps": lr["atom_steps_per_s"] / rate_eq,
                    "variance_gain_measured": P_eq / lr["P"]})
        log(f"equal eps_F: analytic needs P = {P_eq:.1f}; learned speedup "
            f"{out['speedup_at_equal_eps']:.3f}x")
    return out, learned_state, prop, step


def batch_eval(wl, args):
    """c5: train one proposal network, tabulate it on every replica's own k-lattice,
    then time the batched step against the same replicas run one after another."""
    import torch

    from copy_gpu_amd import ForceConfig, systems
    from copy_gpu_amd import proposal as PR
    from copy_gpu_amd.batch import ReplicaBatch

    sizes = systems.replica_sizes(wl["replicas"], wl["n_min"], wl["n_max"], seed=SEED)
    log(f"replicas: {len(sizes)}, total {sum(sizes)} ions, sizes {min(sizes)}..{max(sizes)}")
    alpha, P = wl["alpha"], wl["P"]
    t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to("cuda", torch.float64)  # noqa
    syss = []
    for r, n in enumerate(sizes):
        x, y, z, q, L = systems.random_ions(n, wl["rho"], 0.8, seed=SEED + 1000 + r)
        vx, vy, vz = systems.velocities(n, 1.0, seed=SEED + 1000 + r)
        syss.append((n, L, t64(q), State(*(t64(a) for a in (x, y, z, vx, vy, vz)),
                                          [torch.zeros(n, device="cuda", dtype=torch.float32)
                                           for _ in range(3)])))
    # proposal network: trained on FP64 spectra of the largest replica's trajectory
    big = max(range(len(sizes)), key=lambda r: sizes[r])
    n, L, Q, st = syss[big]
    wlb = dict(wl, n=n)
    f = make_force(wlb, L, P)
    f.Init(Q)
    tr = st.clone()
    f.sample_kvectors(P, SEED, 0)
    f.Execute(tr.X, tr.Y, tr.Z, *tr.F)
    tr.run(f, wlb, 100, 1)
    snaps = []
    for i in range(20):
        tr.run(f, wlb, 10, 101 + 10 * i)
        snaps.append((tr.X.clone(), tr.Y.clone(), tr.Z.clone()))
    f.close()
    sig = math.sqrt(alpha / 2.0) * L / math.pi
    Mg = int(math.ceil(3.0 * sig))
    t = time.perf_counter()
    S2 = PR.spectrum(snaps, Q, (L, L, L), alpha, Mg)
    mlp, hist, k_scale = PR.train_proposal(S2, (L, L, L), alpha, Mg)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    del snaps, tr
    log(f"proposal: spectra of 20 snapshots of replica {big} (N={n}, Mg={Mg}) and training "
        f"{t_train:.2f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")

    cfgs = [ForceConfig(box_len=(L, L, L), cut_off=wl["rc_lj"], coulomb_cut=wl["rc_coul"],
                        skin=0.3, alpha=alpha, kspace="RBE", coulomb_sample_num=P,
                        precision=32, n_max=n) for (n, L, _, _) in syss]
    b = ReplicaBatch(cfgs)
    props = []
    for r, ((n, L, Q, st), fr) in enumerate(zip(syss, b.replicas)):
        fr.Init(Q)
        Mr = max(8, int(math.ceil(3.0 * math.sqrt(alpha / 2.0) * L / math.pi)))
        d = PR.grid_density(mlp, (L, L, L), alpha, Mr, k_scale)
        pr = PR.Proposal(Mr, d, 0.9, alpha, (L, L, L))
        pr.attach(fr)
        props.append(pr)
        fr.sample_kvectors(P, SEED + r, 0)
        fr.Execute(st.X, st.Y, st.Z, *st.F)
    seeds = [SEED + r for r in range(len(syss))]
    states = [(s.X, s.Y, s.Z, s.VX, s.VY, s.VZ, *s.F) for (_, _, _, s) in syss]
    K = args.steps if args.steps is not None else wl["steps"]
    use_graph = not args.no_graph
    W = graph_warmup(args.warmup, K) if use_graph else args.warmup
    b.md_run(states, dt=wl["dt"], nsteps=W, seeds=seeds, step0=1, use_graph=use_graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.md_run(states, dt=wl["dt"], nsteps=K, seeds=seeds, step0=1 + W, use_graph=use_graph)
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    tot = sum(sizes)
    rate_b = tot * K / el_b
    log(f"batched: {K} steps of {len(sizes)} replicas in {el_b * 1e3:.1f} ms -> "
        f"{rate_b:.4e} atom-steps/s ({el_b * 1e3 / K:.3f} ms per batch step)")
    # the same replicas, one md_run after another (each with its own graph)
    for r, fr in enumerate(b.replicas):
        s_ = syss[r][3]
        fr.md_run(s_.X, s_.Y, s_.Z, s_.VX, s_.VY, s_.VZ, *s_.F, dt=wl["dt"], nsteps=W,
                  seed=seeds[r], step0=1 + W + K, use_graph=use_graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r, fr in enumerate(b.replicas):
        s_ = syss[r][3]
        fr.md_run(s_.X, s_.Y, s_.Z, s_.VX, s_.VY, s_.VZ, *s_.F, dt=wl["dt"], nsteps=K,
                  seed=seeds[r], step0=1 + 2 * W + K, use_graph=use_graph)
    torch.cuda.synchronize()
    el_s = time.perf_counter() - t0
    rate_s = tot * K / el_s
    log(f"sequential: {rate_s:.4e} atom-steps/s ({el_s * 1e3 / K:.3f} ms per step of all); "
        f"batched / sequential = {rate_b / rate_s:.2f}x")
    # force error vs FP64 Ewald on four replicas (smallest, two middle, largest)
    order = sorted(range(len(sizes)), key=lambda r: sizes[r])
    pick = [order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]]
    errs = []
    if not args.skip_ewald:
        for r in pick:
            n, L, Q, s_ = syss[r]
            wlr = dict(wl, n=n)
            refs, _ = ewald_refs(wlr, L, Q, [(s_.X, s_.Y, s_.Z)])
            e_l, ek_l = force_error(b.replicas[r], P, [(s_.X, s_.Y, s_.Z)], refs, 10)
            errs.append({"replica": r, "n": n, "eps_F": e_l, "eps_F_kspace_part": ek_l})
            log(f"replica {r} (N={n}): eps_F {e_l:.4e} (k-space part {ek_l:.4e})")
    b.close()
    return {"value": rate_b, "seconds": el_b, "steps": K, "total_atoms": tot,
            "sequential_atom_steps_per_s": rate_s, "batched_over_sequential": rate_b / rate_s,
            "ms_per_batch_step": el_b * 1e3 / K, "ms_per_sequential_sweep": el_s * 1e3 / K,
            "replica_sizes": sizes, "eps_F_samples": errs, "proposal_train_s": t_train}


def make_state(wl, dev):
    """Starting configuration of a workload: positions, charges, velocities."""
    import torch

    from copy_gpu_amd import systems

    n = wl["n"]
    if wl.get("fcc"):
        x, y, z, L = systems.fcc(wl["fcc"], wl["rho"])
        q = np.zeros(n)
    else:
        x, y, z, q, L = systems.random_ions(n, wl["rho"], 0.8, seed=SEED)
    vx, vy, vz = systems.velocities(n, 1.0, seed=SEED)
    t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(dev, torch.float64)  # noqa
    fdt = torch.float64 if wl.get("precision", 32) == 64 else torch.float32
    st = State(*(t64(a) for a in (x, y, z, vx, vy, vz)),
               [torch.zeros(n, device=dev, dtype=fdt) for _ in range(3)])
    return st, t64(q), L, (x, y, z, q, vx, vy, vz)


def roofline_of(f, st, wl, wname, K, step, measured_limiter):
    """Pair-pass roofline of a handle's current state plus per-kernel ms per step."""
    import torch

    per_step_ms, pair_bytes, pair_ms, E, step = kernel_times(f, st, wl, min(K, 100), step)
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    valu = read_traffic(f"pair_{wname}", "valu_wave_instructions_per_launch")
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    traffic = read_traffic(f"pair_{wname}")
    roof = {
        "kernel": "pair",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "traffic_source": (f"profiles/pair_{wname}_traffic.txt (rocprofv3 --pmc FETCH_SIZE and "
                           "WRITE_SIZE of the same pair pass, scaled by the counters' reading "
                           "of a copy kernel of known size)") if traffic else None,
        "algorithmic_bytes_per_launch": pair_bytes,
        "mean_launch_ms": pair_ms,
        "list_entries": E,
        "measured_limiter": measured_limiter,
        "valu_issue_frac": (valu / (pair_ms * 1e-3 * n_cu * ENGINE_CLOCK_HZ)) if valu else None,
        "valu_issue_note": "SQ_INSTS_VALU per launch (profiles/pair_<workload>_traffic.txt) / "
                           "(mean launch time x CUs x 2.4 GHz x 1 wave64 instruction per "
                           "CU-cycle); an indicator of VALU issue, not a bound",
    }
    return roof, per_step_ms, max(per_step_ms, key=per_step_ms.get), step


PAIR_LIMITER = ("VALU issue, LDS gathers and latency together (the bytes moved equal the "
                "algorithmic bytes, so the pass does not reach the HBM roofline; at 100,000 "
                "ions its waves spend 49% of their time waiting on LDS or memory counters, "
                "23% ready but not issued and 28% issuing, with the VALU busy about 75% of "
                "the launch: profiles/pair_c4_pmc.md)")


def plain_leg(wname, args, W, K, use_graph, with_nvt=False):
    """A workload with the analytic sampler (or no k-space): timed MD steps, the
    pair-pass roofline, eps_F against FP64 Ewald on the final configuration."""
    import torch

    wl = WORKLOADS[wname]
    n, P = wl["n"], wl["P"]
    dev = torch.device("cuda:0")
    precision = wl.get("precision", 32)
    kspace = wl.get("kspace", "RBE")
    log(f"workload {wname}: {n} atoms")
    st, Q, L, host = make_state(wl, dev)
    f = make_force(wl, L, P, kspace=kspace, precision=precision)
    f.Init(Q)
    if kspace == "RBE":
        f.sample_kvectors(P, SEED, 0)
    f.Execute(st.X, st.Y, st.Z, *st.F)
    if args.tune:
        t = time.perf_counter()
        shapes = f.autotune(st.X, st.Y, st.Z, *st.F)
        log(f"autotune ({time.perf_counter() - t:.2f} s): {shapes}")
    if args.thermostat:
        f.set_thermostat(args.thermostat, (1.0, 1.0, 0.5))
    log(f"warmup {graph_warmup(W, K) if use_graph else W} steps, timed {K} steps "
        f"(graph={use_graph}, "
        f"thermostat={args.thermostat or 'none'})")
    el, rebuilds, step = timed(f, st, wl, W, K, 1, use_graph)
    value = n * K / el
    log(f"timed {K} steps: {el * 1e3:.2f} ms -> {value:.4e} atom-steps/s, "
        f"{rebuilds} list builds")
    log(f"list/tile stats: {f.stats()}")
    roof, per_step_ms, dominant, step = roofline_of(f, st, wl, wname, K, step, PAIR_LIMITER)
    eps_total = eps_k = None
    if not args.skip_ewald and wl.get("ewald", True):
        log("FP64 Ewald reference on the final configuration")
        refs, nkv = ewald_refs(wl, L, Q, [(st.X, st.Y, st.Z)])
        eps_total, eps_k = force_error(f, P, [(st.X, st.Y, st.Z)], refs, 20)
        log(f"force RMS error vs FP64 Ewald: total {eps_total:.3e}, relative to k-space "
            f"part {eps_k:.3e} ({nkv} half-space k-vectors)")
    out = {"workload": wl["desc"], "value": value, "unit": "atom-steps/s",
           "ms_per_step": el * 1e3 / K, "steps": K, "list_builds_in_timed_steps": rebuilds,
           "force_rms_err_vs_fp64_ewald": eps_total,
           "force_rms_err_vs_fp64_ewald_kspace_part": eps_k,
           "roofline": roof, "per_kernel_ms_per_step": per_step_ms,
           "dominant_kernel": dominant}
    if with_nvt:
        log("NVT (Berendsen) run of the same steps")
        out["nvt_berendsen"] = nvt_leg(f, st, wl, W, K, step, use_graph)
        log(f"NVT: {out['nvt_berendsen']['value']:.4e} atom-steps/s, "
            f"{out['nvt_berendsen']['ms_per_step'] * 1e3:.1f} us per step")
    return out, f, st, host, L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-ewald", action="store_true")
    ap.add_argument("--skip-scale", action="store_true",
                    help="skip the legs reported beside the headline: the 32,768-ion "
                         "workload (timed steps, roofline, Berendsen NVT) and the pair-pass "
                         "roofline of the 1,048,576-ion workload")
    ap.add_argument("--pair-only", action="store_true",
                    help="only the pair-pass roofline of the workload's starting "
                         "configuration (20 force calls; for counter collection)")
    ap.add_argument("--thermostat", default=None,
                    choices=["BERENDSEN", "TEMP_RESCALE", "NOSE_HOOVER", "LANGEVIN"],
                    help="time the MD steps under this NVT thermostat (temperature 1.0, "
                         "Tdamp 0.5) instead of plain velocity Verlet")
    ap.add_argument("--tune", action="store_true",
                    help="run the library's tile-shape autotuner on the starting configuration "
                         "before the timed steps")
    ap.add_argument("--half-shell", action="store_true",
                    help="FP32 handles use the opt-in half-shell pair pass (each pair "
                         "evaluated once, Newton's third law) instead of full rows")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB three times with the library's plain copy kernel "
                         "(known HBM traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    global HALF_SHELL
    HALF_SHELL = args.half_shell
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the passes "
              "are not decomposed over several GPUs", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        import ctypes

        from copy_gpu_amd import _lib

        nbytes = 256 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            rc = _lib.lib().rbmd_util_copy(ctypes.c_void_p(dst.data_ptr()),
                                          ctypes.c_void_p(src.data_ptr()), nbytes, None)
            assert rc == 0
        torch.cuda.synchronize()
        log("copy calibration: 3 x 256 MiB copied")

    if args.pair_only:
        print(json.dumps({"pair_pass": pair_pass_at_scale(args.workload)}), flush=True)
        return

    wl = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else wl["steps"]
    W = args.warmup
    use_graph = not args.no_graph
    base = {"metric": METRIC, "unit": "atom-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic"}
    if wl.get("batch"):
        res = batch_eval(wl, args)
        out = dict(base, value=res["value"], ms_per_step=res["ms_per_batch_step"],
                   dtype="fp32",
                   config={"workload": wl["desc"], "n_atoms_total": res["total_atoms"],
                           "replicas": len(res["replica_sizes"]), "P": wl["P"],
                           "kspace": "learned MLP proposal per replica (mixture 0.9/0.1)",
                           "hip_graph": use_graph, "parallelism": "single GPU"},
                   batch={k: v for k, v in res.items() if k not in ("value",)},
                   roofline=None, cpu_baseline=None)
        print(json.dumps(out), flush=True)
        return

    dev = torch.device("cuda:0")
    precision = wl.get("precision", 32)
    kspace = wl.get("kspace", "RBE")
    if wl.get("learned"):
        st, Q, L, host = make_state(wl, dev)
        learned, (f, st), prop, step = learned_eval(wl, L, Q, st, 1, args)
        lr = learned["runs"][0]
        value, ms_step, rebuilds = lr["atom_steps_per_s"], lr["ms_per_step"], lr["list_builds"]
        eps_total, eps_k = lr["eps_F"], lr["eps_F_kspace_part"]
        log(f"list/tile stats: {f.stats()}")
        roof, per_step_ms, dominant, step = roofline_of(f, st, wl, args.workload, K, step,
                                                        PAIR_LIMITER)
        leg = None
    else:
        leg, f, st, host, L = plain_leg(args.workload, args, W, K, use_graph,
                                        with_nvt=(args.workload == "c2" and not args.skip_scale
                                                  and not args.thermostat))
        learned = None
        value, ms_step = leg["value"], leg["ms_per_step"]
        rebuilds = leg["list_builds_in_timed_steps"]
        eps_total = leg["force_rms_err_vs_fp64_ewald"]
        eps_k = leg["force_rms_err_vs_fp64_ewald_kspace_part"]
        roof, per_step_ms, dominant = (leg["roofline"], leg["per_kernel_ms_per_step"],
                                       leg["dominant_kernel"])
    n, P = wl["n"], wl["P"]
    cpu = None
    if not args.skip_cpu and wl.get("cpu", True):
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)")
        x, y, z, q, vx, vy, vz = host
        cpu = cpu_baseline(wl, x, y, z, q, vx, vy, vz, L, args.cpu_seconds)
        log(f"CPU baseline: {cpu['value']:.4e} atom-steps/s on {cpu['cores']} threads")
    f.close()
    del st
    torch.cuda.synchronize()

    out = dict(base, value=value, ms_per_step=ms_step,
               dtype="fp64" if precision == 64 else "fp32")
    out["config"] = {
        "workload": wl["desc"],
        "n_atoms": n,
        "kspace": ("learned MLP proposal (mixture 0.9 learned / 0.1 analytic)"
                   if learned else "plain RBE (analytic Gaussian sampler)"
                   if kspace == "RBE" else "none (LJ only)"),
        "P": P,
        "precision": ("FP64 throughout" if precision == 64 else
                      "FP32 pair and k-space maths; FP64 positions, integrator, energy"),
        "hip_graph": use_graph,
        "graph_capture_before_timing": use_graph,
        "warmup_steps_run": graph_warmup(W, K) if use_graph else W,
        "thermostat": args.thermostat,
        "list_builds_in_timed_steps": rebuilds,
        "force_rms_err_vs_fp64_ewald": eps_total,
        "force_rms_err_vs_fp64_ewald_kspace_part": eps_k,
        "parallelism": "single GPU",
    }
    out["roofline"] = roof
    out["per_kernel_ms_per_step"] = per_step_ms
    out["dominant_kernel"] = dominant
    out["cpu_baseline"] = cpu
    if learned:
        out["learned_vs_plain"] = learned
    if leg and "nvt_berendsen" in leg:
        out["nvt_berendsen"] = leg["nvt_berendsen"]
    if args.workload == "c4" and not args.skip_scale:
        # beside the headline: the 32,768-ion plain-RBE workload (with its Berendsen NVT
        # leg) and the pair pass of the 1,048,576-ion workload, where the pass is long
        # enough that launch and staging latency no longer dominate it
        c2, fc, _, _, _ = plain_leg("c2", args, W, WORKLOADS["c2"]["steps"], use_graph,
                                    with_nvt=True)
        fc.close()
        out["c2_plain_rbe"] = c2
        out["roofline_c3_pair_pass"] = pair_pass_at_scale("c3")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
