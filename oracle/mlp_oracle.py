"""CPU oracle of the warm-start network (numpy, FP64) — TEST INFRASTRUCTURE ONLY.

The same forward and backward passes as the device kernels (olf-gpu_amd/csrc/mlp.hip): a 3-layer MLP with GELU,
an embedding row added in the first layer, mean squared error on normalised targets. With ``emulate_bf16`` the
matrix operands are rounded to BF16 (round to nearest even, from FP32) where the device rounds them, so the device
gradients can be compared tightly; without it the passes are plain FP64, which the finite-difference test pins.
The reference has no learned warm start, so nothing here restates reference code.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

PARAMS = ("Win", "Emb", "b1", "W2", "b2", "W3", "b3")


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def inputs(x_base, x_mean, x_std, feat, emulate_bf16=True):
    xn = ((np.asarray(x_base, np.float64) - x_mean) / x_std).astype(np.float32).astype(np.float64)
    X = np.concatenate([xn, np.asarray(feat, np.float64).reshape(len(xn), -1)], axis=1)
    return _r(X, emulate_bf16)


def forward(p, X, branch, emulate_bf16=True):
    e = emulate_bf16
    c1 = X @ _r(p["Win"], e).T
    z1 = c1 + p["b1"] + p["Emb"][branch]
    h1 = _r(gelu(z1), e)
    z2 = h1 @ _r(p["W2"], e).T + p["b2"]
    h2 = _r(gelu(z2), e)
    y = h2 @ _r(p["W3"], e).T + p["b3"]
    return dict(z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def loss_and_grads(p, X, branch, target, emulate_bf16=True):
    e = emulate_bf16
    a = forward(p, X, branch, e)
    B, N = a["y"].shape
    d = a["y"] - target
    loss = float((d * d).mean())
    dy = 2.0 * d / (B * N)
    dyr = _r(dy, e)
    g = {"W3": dyr.T @ a["h2"], "b3": dy.sum(0)}
    dz2 = (dyr @ _r(p["W3"], e)) * gelu_grad(a["z2"])
    dz2r = _r(dz2, e)
    g["W2"] = dz2r.T @ a["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (dz2r @ _r(p["W2"], e)) * gelu_grad(a["z1"])
    g["Win"] = _r(dz1, e).T @ X
    g["b1"] = dz1.sum(0)
    gE = np.zeros_like(p["Emb"])
    np.add.at(gE, branch, dz1)
    g["Emb"] = gE
    return loss, g
