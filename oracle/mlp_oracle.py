"""CPU oracle of the learned proposal's network -- test infrastructure only.

Float64 NumPy restatement of the 4 -> 64 -> 64 -> 1 ReLU MLP the engine's
mlp.hip kernels implement: forward, mean-squared-error loss and its
gradient by explicit back-propagation.  Weights are rounded to BF16 first
(the device computes with BF16 weights and FP32 accumulation).  With
bf16_products (the default, as on the device) the inputs of the three
64 x 64 products that run on the matrix cores are rounded to BF16 too: h1
entering layer 2 (z2 = bf16(h1) W2^T) and dz2 entering the backward
products (dh1 = bf16(dz2) W2, gW2 = bf16(dz2)^T bf16(h1)); the other
gradients use the unrounded values.  bf16_products=False is the plain
float64 network (its gradient is checked by finite differences).
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


def _rnd(a: np.ndarray, on: bool) -> np.ndarray:
    return to_bf16(a).astype(np.float64) if on else a


def forward(p, x, bf16_products: bool = True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = _rnd(h1, bf16_products) @ W2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def loss_and_grad(p, x, y, bf16_products: bool = True):
    """MSE over the rows of x and its gradient w.r.t. the packed parameters."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    m = x.shape[0]
    s, (z1, h1, z2, h2) = forward(p, x, bf16_products)
    r = s - y
    loss = float((r * r).mean())
    ds = 2.0 * r / m
    g = np.zeros(N_PARAMS)
    g[O_B3] = ds.sum()
    g[O_W3:O_B3] = h2.T @ ds
    dz2 = np.outer(ds, w3) * (z2 > 0)
    g[O_B2:O_W3] = dz2.sum(0)
    dz2r = _rnd(dz2, bf16_products)
    g[O_W2:O_B2] = (dz2r.T @ _rnd(h1, bf16_products)).ravel()
    dz1 = (dz2r @ W2) * (z1 > 0)
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W1:O_B1] = (dz1.T @ x).ravel()
    return loss, g
