"""CPU oracle of the learned proposal's network -- test infrastructure only.

Float64 NumPy restatement of the 4 -> 64 -> 64 -> 1 ReLU MLP the engine's
mlp.hip kernels implement: forward, mean-squared-error loss and its
gradient by explicit back-propagation, and one Adam step.  Weights are
rounded to BF16 first (the device computes with BF16 weights and FP32
accumulation).
"""
from __future__ import annotations

import numpy as np

K_IN, K_H = 4, 64
O_W1 = 0
O_B1 = O_W1 + K_H * K_IN
O_W2 = O_B1 + K_H
O_B2 = O_W2 + K_H * K_H
O_W3 = O_B2 + K_H
O_B3 = O_W3 + K_H
N_PARAMS = O_B3 + 1


def to_bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (round to nearest even) and back."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(K_H, K_IN), p[O_B1:O_W2], p[O_W2:O_B2].reshape(K_H, K_H), p[O_B2:O_W3],
            p[O_W3:O_B3], p[O_B3])


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def loss_and_grad(p, x, y):
    """MSE over the rows of x and its gradient w.r.t. the packed parameters."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    m = x.shape[0]
    s, (z1, h1, z2, h2) = forward(p, x)
    r = s - y
    loss = float((r * r).mean())
    ds = 2.0 * r / m
    g = np.zeros(N_PARAMS)
    g[O_B3] = ds.sum()
    g[O_W3:O_B3] = h2.T @ ds
    dz2 = np.outer(ds, w3) * (z2 > 0)
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W2:O_B2] = (dz2.T @ h1).ravel()
    dz1 = (dz2 @ W2) * (z1 > 0)
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W1:O_B1] = (dz1.T @ x).ravel()
    return loss, g
