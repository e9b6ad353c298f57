#!/usr/bin/env python3
"""Benchmark: atom-steps/s of the charged-LJ + Random Batch Ewald MD step on one MI355X.

Workload (default): 32,768 +/-1 ions, uniform random with minimum separation
0.8, density 0.3 (L = 47.80), LJ cutoff 2.5, erfc real-space cutoff 5.0,
alpha 0.5, P = 100 k-vectors resampled every step from the RBE Gaussian,
Verlet skin 0.3, T = 1, dt = 0.002; FP32 pair and k-space maths, FP64
positions, velocities, integrator.  A step = kick, drift, list check (and
rebuild when needed), k-sampling, structure factor, pair pass, k-space force,
kick -- captured once in a HIP graph and replayed.

Prints ONE JSON line on stdout (logs go to stderr).  Also reported:
  * force RMS error vs FP64 full Ewald on the final configuration, over 20
    independent k-lists (total force and k-space part);
  * roofline of the pair pass (dominant HBM-bound kernel): algorithmic bytes
    4 E + 40 N per launch over the mean launch time from HIP events on its
    stream (E = full-list entries at r_list), against 8 TB/s;
  * the CPU baseline: the reference-faithful FP32 step of the oracle on the
    host cores for a bounded sample of steps.
"""
from __future__ import annotations

import argparse
import json
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
    "c2": dict(n=32768, rho=0.3, rc_lj=2.5, rc_coul=5.0, alpha=0.5, P=100, dt=0.002,
               steps=500, desc="charged LJ electrolyte, plain RBE: 32768 ions, density 0.3, "
                           "rc_LJ 2.5, rc_coul 5.0, alpha 0.5, P=100, dt 0.002"),
    "c3": dict(n=1048576, rho=0.3, rc_lj=2.5, rc_coul=6.0, alpha=0.3, P=500, dt=0.002,
               steps=200, desc="large charged LJ, plain RBE: 1048576 ions, density 0.3, "
                           "rc_LJ 2.5, rc_coul 6.0, alpha 0.3, P=500, dt 0.002"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 --pmc summary
    (profiles/<name>_traffic.txt, line 'bytes_per_launch: <value>'), else None."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(wl, x, y, z, q, vx, vy, vz, L, target_s):
    import oracle

    n, P = wl["n"], wl["P"]
    threads = max(1, min(16, os.cpu_count() or 1))
    pos = np.ascontiguousarray(np.stack([x, y, z], 1), np.float32)
    vel = np.ascontiguousarray(np.stack([vx, vy, vz], 1), np.float32)
    frc = np.zeros_like(pos)
    from copy_gpu_amd import systems

    def klist(steps, s0):
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
    return {"value": n * steps / el, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} reference-faithful FP32 steps (per-step cell sort, list at rc "
                      f"without skin, N*P structure-factor arrays) of the same {n}-ion "
                      f"workload, OpenMP over atoms, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-ewald", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB three times with the library's plain copy kernel "
                         "(known HBM traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU "
              "decomposition is not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch

    from copy_gpu_amd import ForceConfig, LJCutCoulKspaceForce, systems

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

    wl = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else wl["steps"]
    W = args.warmup
    n, P, alpha = wl["n"], wl["P"], wl["alpha"]
    dev = torch.device("cuda:0")
    log(f"workload {args.workload}: generating {n} ions")
    x, y, z, q, L = systems.random_ions(n, wl["rho"], 0.8, seed=SEED)
    vx, vy, vz = systems.velocities(n, 1.0, seed=SEED)
    t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(dev, torch.float64)  # noqa
    cfg = ForceConfig(box_len=(L, L, L), cut_off=wl["rc_lj"], coulomb_cut=wl["rc_coul"],
                      skin=0.3, alpha=alpha, kspace="RBE", coulomb_sample_num=P,
                      precision=32, n_max=n)
    f = LJCutCoulKspaceForce(cfg)
    f.Init(t64(q))
    X, Y, Z, VX, VY, VZ = (t64(a) for a in (x, y, z, vx, vy, vz))
    F = [torch.zeros(n, device=dev, dtype=torch.float32) for _ in range(3)]
    f.sample_kvectors(P, SEED, 0)
    f.Execute(X, Y, Z, *F)
    use_graph = not args.no_graph
    step = 1
    log(f"warmup {W} steps (graph={use_graph})")
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=wl["dt"], nsteps=W, seed=SEED, step0=step,
             use_graph=use_graph)
    step += W
    torch.cuda.synchronize()
    b0 = f.stats()["builds"]
    t0 = time.perf_counter()
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=wl["dt"], nsteps=K, seed=SEED, step0=step,
             use_graph=use_graph)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    step += K
    st = f.stats()
    rebuilds = st["builds"] - b0
    value = n * K / el
    log(f"timed {K} steps: {el * 1e3:.2f} ms -> {value:.4e} atom-steps/s, {rebuilds} list builds")
    log(f"list/tile stats: {st}")

    # per-kernel device times from HIP events (eager launches of the same step)
    f.set_timing(True)
    Kt = min(K, 100)
    f.md_run(X, Y, Z, VX, VY, VZ, *F, dt=wl["dt"], nsteps=Kt, seed=SEED, step0=step,
             use_graph=False)
    step += Kt
    torch.cuda.synchronize()
    names = ["check", "bin", "scan", "scatter", "cellsort", "list", "gather", "sample", "rho",
             "rho_final", "pair", "kforce", "kick"]
    per = {}
    for nm in names:
        ms, cnt = f.timing(nm)
        per[nm] = {"mean_ms": ms, "launches": cnt}
    f.set_timing(False)
    st = f.stats()
    E = st["entries"]
    pair_bytes = 4.0 * E + 40.0 * n
    pair_ms = per["pair"]["mean_ms"]
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    per_step_ms = {k: v["mean_ms"] * v["launches"] / Kt for k, v in per.items()}
    dominant = max(per_step_ms, key=per_step_ms.get)
    log("per-step device ms: " + ", ".join(f"{k} {v:.4f}" for k, v in per_step_ms.items()))

    # force RMS error vs FP64 full Ewald on the final configuration
    eps_total = eps_k = None
    if not args.skip_ewald:
        log("FP64 Ewald reference on the final configuration")
        ew = LJCutCoulKspaceForce(ForceConfig(box_len=(L, L, L), cut_off=wl["rc_lj"],
                                              coulomb_cut=wl["rc_coul"], skin=0.3, alpha=alpha,
                                              kspace="EWALD", precision=64, n_max=n,
                                              ewald_tol=1e-14))
        ew.Init(t64(q))
        Fe = [torch.zeros(n, device=dev, dtype=torch.float64) for _ in range(3)]
        ew.Execute(X, Y, Z, *Fe)
        nk = LJCutCoulKspaceForce(ForceConfig(box_len=(L, L, L), cut_off=wl["rc_lj"],
                                              coulomb_cut=wl["rc_coul"], skin=0.3, alpha=alpha,
                                              kspace="NONE", precision=64, n_max=n))
        nk.Init(t64(q))
        Fr = [torch.zeros(n, device=dev, dtype=torch.float64) for _ in range(3)]
        nk.Execute(X, Y, Z, *Fr)
        Fe_t = torch.stack(Fe, 1)
        Fek = Fe_t - torch.stack(Fr, 1)
        # the real-space and LJ parts are identical, so the same difference
        # is normalised by the total force and by the Ewald k-space force
        num = 0.0
        for s in range(20):
            f.sample_kvectors(P, SEED + 1, s)
            f.Execute(X, Y, Z, *F)
            Fg = torch.stack(F, 1).double()
            num += float(((Fg - Fe_t) ** 2).sum())
        eps_total = (num / (float((Fe_t ** 2).sum()) * 20)) ** 0.5
        eps_k = (num / (float((Fek ** 2).sum()) * 20)) ** 0.5
        log(f"force RMS error vs FP64 Ewald: total {eps_total:.3e}, relative to k-space part "
            f"{eps_k:.3e} ({ew.stats()['ewald_k']} half-space k-vectors)")
        ew.close()
        nk.close()

    cpu = None
    if not args.skip_cpu:
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(wl, x, y, z, q, vx, vy, vz, L, args.cpu_seconds)
        log(f"CPU baseline: {cpu['value']:.4e} atom-steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": el * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": wl["desc"],
            "n_atoms": n,
            "kspace": "plain RBE (analytic Gaussian sampler)",
            "P": P,
            "precision": "FP32 pair and k-space maths; FP64 positions, integrator, energy",
            "hip_graph": use_graph,
            "list_builds_in_timed_steps": rebuilds,
            "force_rms_err_vs_fp64_ewald": eps_total,
            "force_rms_err_vs_fp64_ewald_kspace_part": eps_k,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "pair",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(f"pair_{args.workload}"),
            "algorithmic_bytes_per_launch": pair_bytes,
            "mean_launch_ms": pair_ms,
            "list_entries": E,
        },
        "per_kernel_ms_per_step": per_step_ms,
        "dominant_kernel": dominant,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
