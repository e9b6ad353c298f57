"""Pins of the FP64 proposal-network oracle (oracle/mlp.py).

The network has no counterpart in the reference, so the oracle is pinned by
mathematics: central finite differences of the loss for every parameter,
closed forms of special weight settings, and torch's own float32 -> bfloat16
rounding for the BF16 weight copy.
"""
import numpy as np
import pytest
import torch

from oracle import mlp


def _rand_params(rng, scale=0.5):
    p = rng.normal(0, scale, mlp.NPARAM)
    return p


def test_layout_size():
    assert mlp.NPARAM == 4545
    W1, b1, W2, b2, w3, b3 = mlp.unpack(np.arange(mlp.NPARAM, dtype=np.float64))
    assert W1.shape == (64, 4) and W2.shape == (64, 64)
    assert b1[0] == 256 and W2[0, 0] == 320 and b2[0] == 4416 and w3[0] == 4480 and b3 == 4544


def test_gradient_finite_differences():
    rng = np.random.default_rng(7)
    p = _rand_params(rng, 0.3)
    f = rng.uniform(0, 2, (12, 4))
    y = rng.normal(0, 1, 12)
    w = rng.uniform(0.1, 1, 12)
    loss, g = mlp.loss_grad(p, f, y, w)
    h = 1e-6
    fd = np.empty(mlp.NPARAM)
    for i in range(mlp.NPARAM):
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        fd[i] = (mlp.loss_grad(p + e, f, y, w)[0] - mlp.loss_grad(p - e, f, y, w)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-6 * max(1.0, np.abs(g).max())
    # a dropped term (e.g. the SiLU derivative) would fail by far more than this
    assert np.abs(g).max() > 1e-3


def test_forward_closed_forms():
    p = np.zeros(mlp.NPARAM)
    f = np.array([[0.3, 0.1, 0.7, 0.9]])
    # all zero: silu(0) = 0 -> z = b3 = 0 -> h = ln 2
    assert mlp.forward(p, f)[0][0] == pytest.approx(np.log(2.0), rel=1e-15)
    # one path: a1_0 = silu(w . f), a2_0 = silu(W2[0,0] a1_0), z = w3_0 a2_0 + b3
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p)
    q = p.copy()
    q[0:4] = [1.0, -2.0, 0.5, 0.25]
    q[320] = 1.5          # W2[0,0]
    q[4480] = 2.0         # w3[0]
    q[4544] = -0.3        # b3
    pre = 1.0 * 0.3 - 2.0 * 0.1 + 0.5 * 0.7 + 0.25 * 0.9
    a1 = pre / (1 + np.exp(-pre))
    a2 = 1.5 * a1 / (1 + np.exp(-1.5 * a1))
    z = 2.0 * a2 - 0.3
    assert mlp.forward(q, f)[0][0] == pytest.approx(np.log1p(np.exp(z)), rel=1e-14)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 500),
                        [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16_round(x), ref)


def test_features():
    m = np.array([[1, -2, 0], [0, 0, 3]])
    f = mlp.features(m, (10.0, 20.0, 30.0), 2.0)
    k = 2 * np.pi * np.array([[0.1, -0.1, 0.0], [0, 0, 0.1]])
    assert np.allclose(f[:, :3], np.abs(k) / 2.0, rtol=1e-15)
    assert np.allclose(f[:, 3], np.linalg.norm(k, axis=1) / 2.0, rtol=1e-15)
