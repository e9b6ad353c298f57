"""Learned k-space importance proposal for Random Batch Ewald.

The reference draws the RBE k-vectors from the analytic Gaussian
q(m) ~ a(m) = exp(-k^2 / 4 alpha) (RBEPSAMPLE, rbmd/framework/model/RBEPSample.h:143-175,
called from ExecutionNVT::PreForce, rbmd/framework/run/ExecutionNVT.cpp:331-353).
Here a small MLP h(k) reshapes it: q(m) ~ a(m) h(k), mixed with the analytic
sampler (weight 1 - lambda), and each sampled k carries w = a(m) / (P q(m)) so the
force and energy estimators stay unbiased.

Why h(k) ~ sqrt(<|rho(k)|^2>) / |k|: the RBE force of one sample on atom i is
w f_i(k) with sum_i |f_i(k)|^2 ~ (4 pi / V)^2 N |rho(k)|^2 / (2 k^2), so the
atom-summed force variance sum_k a(k)^2 |f(k)|^2 / q(k) is smallest for
q ~ a(k) |rho(k)| / |k| (Cauchy-Schwarz).  The network is trained in-project on
<|rho(k)|^2> from FP64 structure factors of MD snapshots on the lattice box
|m_d| <= Mg, by its own forward/backward/Adam kernels (csrc/mlp.hip).

Everything runs in librbmd_force.so; numpy only prepares host tables.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from .force import ForceConfig, LJCutCoulKspaceForce, RBMDError, _stream

NPARAM = 4545


def lattice_grid(Mg: int) -> np.ndarray:
    """All m with |m_d| <= Mg, x fastest: int32 [(2Mg+1)^3, 3]."""
    r = np.arange(-Mg, Mg + 1, dtype=np.int32)
    mz, my, mx = np.meshgrid(r, r, r, indexing="ij")
    return np.stack([mx.ravel(), my.ravel(), mz.ravel()], 1)


def half_space(m: np.ndarray) -> np.ndarray:
    """Mask of the half space m > 0 in (z, y, x) lexicographic order."""
    mx, my, mz = m[:, 0], m[:, 1], m[:, 2]
    return (mz > 0) | ((mz == 0) & (my > 0)) | ((mz == 0) & (my == 0) & (mx > 0))


def kvecs(m: np.ndarray, box_len) -> np.ndarray:
    return 2.0 * np.pi * np.asarray(m, np.float64) / np.asarray(box_len, np.float64)


def features(m: np.ndarray, box_len, k_scale: float) -> np.ndarray:
    """Network input (|kx|, |ky|, |kz|, |k|) / k_scale, float32 [B, 4]."""
    k = kvecs(m, box_len)
    kn = np.sqrt((k * k).sum(1))
    return (np.concatenate([np.abs(k), kn[:, None]], 1) / k_scale).astype(np.float32)


# ---------------------------------------------------------------- network
class ProposalMLP:
    """4 -> 64 -> 64 -> 1 MLP (SiLU, softplus) with device forward/backward/Adam."""

    def __init__(self, seed: int = 1):
        self.lib = L.lib()
        h = C.c_void_p()
        rc = self.lib.rbmd_mlp_create(seed, C.byref(h))
        if rc != 0:
            raise RBMDError(f"rbmd_mlp_create -> {rc}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.rbmd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RBMDError(f"{what} -> {rc}: {self.lib.rbmd_mlp_last_error(self.h).decode()}")

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, np.float32)
        self._check(self.lib.rbmd_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p)),
                    "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == NPARAM
        self._check(self.lib.rbmd_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p)),
                    "set_params")

    @staticmethod
    def _dev(a, dtype=torch.float32):
        t = torch.as_tensor(a, dtype=dtype, device="cuda")
        return t.contiguous()

    def forward(self, feat) -> torch.Tensor:
        f = self._dev(feat)
        B = f.shape[0]
        out = torch.empty(B, dtype=torch.float32, device="cuda")
        self._check(self.lib.rbmd_mlp_forward(self.h, C.c_void_p(f.data_ptr()), B,
                                              C.c_void_p(out.data_ptr()), _stream(None)),
                    "forward")
        return out

    def loss_grad(self, feat, y, w):
        f, yt, wt = self._dev(feat), self._dev(y), self._dev(w)
        loss = C.c_double()
        g = np.zeros(NPARAM, np.float32)
        self._check(self.lib.rbmd_mlp_loss_grad(
            self.h, C.c_void_p(f.data_ptr()), C.c_void_p(yt.data_ptr()),
            C.c_void_p(wt.data_ptr()), f.shape[0], C.byref(loss),
            g.ctypes.data_as(C.c_void_p), _stream(None)), "loss_grad")
        return loss.value, g

    def train(self, feat, y, w, iters: int, lr: float) -> float:
        f, yt, wt = self._dev(feat), self._dev(y), self._dev(w)
        loss = C.c_double()
        self._check(self.lib.rbmd_mlp_train(
            self.h, C.c_void_p(f.data_ptr()), C.c_void_p(yt.data_ptr()),
            C.c_void_p(wt.data_ptr()), f.shape[0], iters, lr, C.byref(loss), _stream(None)),
            "train")
        return loss.value


# ---------------------------------------------------------------- tables
class Proposal:
    """Host tables of a mixture proposal (density on the lattice box, weight lambda)."""

    def __init__(self, Mg: int, density: np.ndarray, lam: float, alpha: float, box_len):
        self.lib = L.lib()
        d = np.ascontiguousarray(density, np.float64)
        assert d.size == (2 * Mg + 1) ** 3
        self._L = (C.c_double * 3)(*[float(v) for v in box_len])
        h = C.c_void_p()
        rc = self.lib.rbmd_proposal_create(Mg, d.ctypes.data_as(C.c_void_p), lam, alpha,
                                           self._L, C.byref(h))
        if rc != 0:
            raise RBMDError(f"rbmd_proposal_create -> {rc}")
        self.h = h
        self.Mg, self.lam, self.alpha = Mg, lam, alpha
        self.box_len = tuple(float(v) for v in box_len)
        self.density = d

    def close(self):
        if getattr(self, "h", None):
            self.lib.rbmd_proposal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample_host(self, P: int, seed: int, step: int):
        m = np.zeros((P, 3), np.int32)
        w = np.zeros(P)
        rc = self.lib.rbmd_proposal_sample_host(self.h, P, seed, step,
                                                m.ctypes.data_as(C.c_void_p),
                                                w.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RBMDError(f"rbmd_proposal_sample_host -> {rc}")
        return m, w

    def attach(self, force: LJCutCoulKspaceForce):
        force._check(self.lib.rbmd_force_set_proposal(force.h, self.h), "set_proposal")

    @staticmethod
    def detach(force: LJCutCoulKspaceForce):
        force._check(L.lib().rbmd_force_set_proposal(force.h, None), "set_proposal")


def analytic_weight(m: np.ndarray, box_len, alpha: float) -> np.ndarray:
    k = kvecs(m, box_len)
    return np.exp(-(k * k).sum(1) / (4.0 * alpha))


# ---------------------------------------------------------------- spectra
def spectrum(snapshots, q: torch.Tensor, box_len, alpha: float, Mg: int,
             box_lo=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Mean |rho(m)|^2 over snapshots on the lattice box |m_d| <= Mg (FP64).

    snapshots: iterable of (x, y, z) FP64 device tensors.  rho(m) is the FP64
    structure factor of the engine's k-space pass over the half space; the
    other half follows from rho(-m) = conj(rho(m)).  Returns [(2Mg+1)^3], 0 at m = 0.
    """
    m = lattice_grid(Mg)
    hs = half_space(m)
    mh = torch.as_tensor(m[hs].ravel(), dtype=torch.int32, device="cuda")
    Ph = int(hs.sum())
    n = q.numel()
    cfg = ForceConfig(box_len=tuple(box_len), box_lo=tuple(box_lo), cut_off=1.0,
                      coulomb_cut=1.0, skin=0.1, alpha=alpha, kspace="RBE",
                      coulomb_sample_num=Ph, precision=64, n_max=n)
    f = LJCutCoulKspaceForce(cfg)
    try:
        f.Init(q)
        f.set_kvectors(mh)
        fx = torch.empty(n, dtype=torch.float64, device="cuda")
        fy, fz = torch.empty_like(fx), torch.empty_like(fx)
        acc = np.zeros(Ph)
        cnt = 0
        for (x, y, z) in snapshots:
            f.Execute(x, y, z, fx, fy, fz)
            _, _, re, im = f.kspace()
            acc += re * re + im * im
            cnt += 1
    finally:
        f.close()
    S2 = np.zeros(len(m))
    S2[hs] = acc / max(cnt, 1)
    # rho(-m) = conj(rho(m)): index of -m in x-fastest order is G - 1 - i
    G = len(m)
    idx = np.nonzero(hs)[0]
    S2[G - 1 - idx] = S2[idx]
    return S2


def training_set(S2: np.ndarray, box_len, alpha: float, Mg: int, k_scale: float):
    """Octant-folded (features, log-target, loss weight) from a grid spectrum.

    Target t(k) = sqrt(<|rho|^2>) / |k| averaged over the lattice points that
    share |m_d| (they share the network input); y = log t centred on its
    weighted mean; loss weight ~ a(k) t(k) x multiplicity (where the proposal
    puts its samples), normalised to sum 1.
    """
    m = lattice_grid(Mg)
    side = Mg + 1
    am = np.abs(m)
    key = am[:, 0] + side * (am[:, 1] + side * am[:, 2])
    nz = key != 0
    cnt = np.bincount(key[nz], minlength=side ** 3).astype(np.float64)
    s2 = np.bincount(key[nz], weights=S2[nz], minlength=side ** 3)
    keep = cnt > 0
    ko = np.nonzero(keep)[0]
    mo = np.stack([ko % side, (ko // side) % side, ko // (side * side)], 1)
    s2 = s2[keep] / cnt[keep]
    k = kvecs(mo, box_len)
    kn = np.sqrt((k * k).sum(1))
    a = np.exp(-kn * kn / (4.0 * alpha))
    t = np.sqrt(np.maximum(s2, 1e-300)) / kn
    w = a * t * cnt[keep]
    w /= w.sum()
    y = np.log(t)
    y -= (w * y).sum()
    return features(mo, box_len, k_scale), y.astype(np.float32), w.astype(np.float32), mo


def grid_density(mlp: ProposalMLP, box_len, alpha: float, Mg: int, k_scale: float) -> np.ndarray:
    """Unnormalised proposal a(m) h(k) on the lattice box (network inference on the GPU)."""
    m = lattice_grid(Mg)
    h = mlp.forward(features(m, box_len, k_scale)).double().cpu().numpy()
    d = analytic_weight(m, box_len, alpha) * h
    d[(m == 0).all(1)] = 0.0
    return d


def variance_factor(S2: np.ndarray, density: np.ndarray | None, box_len, alpha: float,
                    Mg: int, lam: float = 1.0) -> float:
    """Per-sample atom-summed force variance proxy sum a^2 t^2 / q over the lattice
    box, for q = mixture of the normalised density (weight lam) and the analytic
    sampler (density None: analytic only).  Ratios of this between proposals are
    the expected ratios of their force variances at equal P."""
    m = lattice_grid(Mg)
    a = analytic_weight(m, box_len, alpha)
    k = kvecs(m, box_len)
    k2 = (k * k).sum(1)
    nz = k2 > 0
    qa = a / a[nz].sum()
    if density is None:
        q = qa
    else:
        q = lam * density / density[nz].sum() + (1.0 - lam) * qa
    return float((a[nz] ** 2 * S2[nz] / k2[nz] / q[nz]).sum())


def train_proposal(S2: np.ndarray, box_len, alpha: float, Mg: int, seed: int = 1,
                   schedule=((1e-2, 600), (3e-3, 900), (1e-3, 900))):
    """Train the network on a grid spectrum; returns (mlp, loss history, k_scale)."""
    k_scale = 2.0 * math.sqrt(alpha)
    feat, y, w, _ = training_set(S2, box_len, alpha, Mg, k_scale)
    mlp = ProposalMLP(seed)
    hist = [mlp.loss_grad(feat, y, w)[0]]
    for lr, it in schedule:
        hist.append(mlp.train(feat, y, w, it, lr))
    return mlp, hist, k_scale


def learned_proposal(S2: np.ndarray, box_len, alpha: float, Mg: int, lam: float = 0.9,
                     seed: int = 1, schedule=None):
    """Train on S2 and tabulate: returns (Proposal, mlp, loss history)."""
    kw = {} if schedule is None else {"schedule": schedule}
    mlp, hist, k_scale = train_proposal(S2, box_len, alpha, Mg, seed, **kw)
    d = grid_density(mlp, box_len, alpha, Mg, k_scale)
    return Proposal(Mg, d, lam, alpha, box_len), mlp, hist
