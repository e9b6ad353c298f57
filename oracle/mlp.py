"""FP64 reference of the learned k-space proposal network -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component; this is the checker for the project's
own MLP kernels (forward, loss, backward) and is imported only by tests/.

Network (4 -> 64 -> 64 -> 1): features f = (|kx|, |ky|, |kz|, |k|) / k_scale,
  a1 = silu(W1 f + b1), a2 = silu(W2 a1 + b2), z = w3 . a2 + b3, h = softplus(z).
Parameter vector layout (4545 values): W1[64][4], b1[64], W2[64][64], b2[64],
w3[64], b3[1].  The proposal density on the k-lattice is exp(-k^2/(4 alpha)) h(k).
Loss over a batch: sum_s w_s (log h_s - y_s)^2.
"""
from __future__ import annotations

import numpy as np

H = 64
NPARAM = H * 4 + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * 4].reshape(H, 4); o += H * 4
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def bf16_round(p):
    """Round FP32 values to bfloat16 (nearest even) and back, as the kernels do."""
    a = np.asarray(p, np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def softplus(z):
    return np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20))))


def forward(p, f):
    """h(f) for features f[B, 4]; also the intermediates."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    pre1 = f @ W1.T + b1
    a1 = silu(pre1)
    pre2 = a1 @ W2.T + b2
    a2 = silu(pre2)
    z = a2 @ w3 + b3
    return softplus(z), (pre1, a1, pre2, a2, z)


def loss_grad(p, f, y, w):
    """Loss and its gradient w.r.t. the parameter vector (FP64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h, (pre1, a1, pre2, a2, z) = forward(p, f)
    t = np.log(h) - y
    loss = float(np.sum(w * t * t))
    sig = 1.0 / (1.0 + np.exp(-z))
    dz = 2.0 * w * t * sig / h
    dw3 = a2.T @ dz
    db3 = dz.sum()
    dpre2 = (dz[:, None] * w3[None, :]) * dsilu(pre2)
    dW2 = dpre2.T @ a1
    db2 = dpre2.sum(0)
    dpre1 = (dpre2 @ W2) * dsilu(pre1)
    dW1 = dpre1.T @ f
    db1 = dpre1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return loss, g


def features(m, box_len, k_scale):
    """(|kx|, |ky|, |kz|, |k|) / k_scale for integer lattice vectors m[B, 3]."""
    k = 2.0 * np.pi * np.asarray(m, np.float64) / np.asarray(box_len, np.float64)
    kn = np.sqrt((k * k).sum(1))
    return np.concatenate([np.abs(k), kn[:, None]], 1) / k_scale
