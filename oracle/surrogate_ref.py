"""CPU reference of the net-line surrogate's forward and backward passes.

TEST INFRASTRUCTURE: only tests/ import this module.  It restates the GPU
kernels' arithmetic in numpy: operands of every product rounded to bf16
(round to nearest even) at the same points as the kernels, products
accumulated in float64; with bf16=False everything stays float64 (the exact
gradient the kernels approximate).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and return them as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _q(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(layers, X, use_bf16=True):
    """layers: [(W [N][K padded], b [N])]; X [B][in]. Returns (Y, cache)."""
    B = X.shape[0]
    K0 = layers[0][0].shape[1]
    a = np.zeros((B, K0))
    a[:, :X.shape[1]] = X
    a = _q(a.astype(np.float32), use_bf16)
    cache = [a]
    zs = []
    for l, (W, b) in enumerate(layers):
        z = a @ _q(W, use_bf16).T + b.astype(np.float64)
        if l + 1 < len(layers):
            zs.append(z)
            a = _q(gelu(z).astype(np.float32), use_bf16) if use_bf16 else gelu(z)
            cache.append(a)
        else:
            return z, (cache, zs)
    raise AssertionError


def loss_and_grads(layers, X, T, M, use_bf16=True):
    """Masked mean-square loss and gradients [(dW, db)] in the kernels' order."""
    Y, (acts, zs) = forward(layers, X, use_bf16)
    count = float((M != 0).sum())
    d = (Y - T) * M
    loss = float((d * d).sum() / count)
    g = 2.0 * d / count
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        gq = _q(g.astype(np.float32), use_bf16)
        dW = gq.T @ acts[l]
        db = g.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            W = _q(layers[l][0], use_bf16)
            g = (gq @ W) * gelu_grad(zs[l - 1])
    return loss, grads
