"""Batched uneven replicas: many independent charged-LJ systems advanced together.

Every phase of the velocity-Verlet step is one launch over all replicas (a table
of (replica, workgroup) work items with per-replica arguments in device memory)
and the step is captured once in a HIP graph (rbmd_batch_* in librbmd_force.so).
Each replica is an ordinary force handle, reachable as `batch.replicas[r]`, with
the reference's per-system ensemble loop (ExecutionNVT::Solve,
rbmd/framework/run/ExecutionNVT.cpp:133-159); the reference itself runs one
system per process.
"""
from __future__ import annotations

import ctypes as C

from . import _lib as L
from .force import ForceConfig, LJCutCoulKspaceForce, RBMDError, _stream


class ReplicaBatch:
    def __init__(self, cfgs: list[ForceConfig]):
        self.lib = L.lib()
        self.cfgs = list(cfgs)
        R = len(self.cfgs)
        params = (L.ForceParams * R)()
        self._keep = []
        for r, cfg in enumerate(self.cfgs):
            p, keep = LJCutCoulKspaceForce.params(cfg)
            params[r] = p
            self._keep.append(keep)
        nmax = (C.c_int64 * R)(*[int(c.n_max) for c in self.cfgs])
        h = C.c_void_p()
        rc = self.lib.rbmd_batch_create(R, params, nmax, C.byref(h))
        if rc != 0:
            raise RBMDError(f"rbmd_batch_create -> {rc}: "
                            f"{self.lib.rbmd_batch_last_error(None).decode()}")
        self.h = h
        self.replicas = []
        for r, cfg in enumerate(self.cfgs):
            hr = self.lib.rbmd_batch_replica(self.h, r)
            self.replicas.append(LJCutCoulKspaceForce(cfg, _borrowed=C.c_void_p(hr)))

    def __len__(self):
        return len(self.replicas)

    def close(self):
        if getattr(self, "h", None):
            for f in self.replicas:
                f.h = None
            self.lib.rbmd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _states(states):
        R = len(states)
        arr = (L.ReplicaState * R)()
        for r, stt in enumerate(states):
            t = list(stt) + [None] * (11 - len(stt))
            arr[r] = L.ReplicaState(*[None if v is None else v.data_ptr() for v in t])
        return arr

    def autotune(self, states, reps=3, stream=None):
        """One tile-shape policy for all replicas, timed as batched launches."""
        assert len(states) == len(self.replicas)
        rc = self.lib.rbmd_batch_autotune(self.h, self._states(states), reps, _stream(stream))
        if rc != 0:
            raise RBMDError(f"rbmd_batch_autotune -> {rc}: "
                            f"{self.lib.rbmd_batch_last_error(self.h).decode()}")
        st = self.replicas[0].stats()
        return {k: st[k] for k in ("split", "tile_stage", "rho_atoms_per_workgroup",
                                   "tuned_cost_us")}

    def md_run(self, states, dt, nsteps, seeds, step0=0, use_graph=True, stream=None):
        """states[r] = (x, y, z, vx, vy, vz, fx, fy, fz[, img[, mass]]) device tensors;
        seeds[r] = replica r's k-sampler seed."""
        R = len(self.replicas)
        assert len(states) == R and len(seeds) == R
        arr = self._states(states)
        sd = (C.c_uint64 * R)(*[int(s) for s in seeds])
        rc = self.lib.rbmd_batch_md_run(self.h, arr, dt, nsteps, sd, step0,
                                        1 if use_graph else 0, _stream(stream))
        if rc != 0:
            raise RBMDError(f"rbmd_batch_md_run -> {rc}: "
                            f"{self.lib.rbmd_batch_last_error(self.h).decode()}")
