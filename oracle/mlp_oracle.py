"""FP64 CPU oracle of the energy-surrogate MLP (n_in -> H -> H -> H -> 1, GELU).

TEST INFRASTRUCTURE ONLY: used by tests to check the engine's hand-written
forward/backward kernels. Parameters use the engine's flat layout
(W1 [k_in x H], b1, W2 [H x H], b2, W3 [H x H], b3, w4 [H], b4).
GELU is the exact form 0.5 z (1 + erf(z / sqrt 2)); the loss is the mean
squared error.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def split(params: np.ndarray, k_in: int, H: int):
    sizes = [k_in * H, H, H * H, H, H * H, H, H, 1]
    out, o = [], 0
    for s in sizes:
        out.append(params[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3, w4, b4 = out
    return W1.reshape(k_in, H), b1, W2.reshape(H, H), b2, W3.reshape(H, H), b3, w4, b4


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, x, k_in, H):
    W1, b1, W2, b2, W3, b3, w4, b4 = split(params, k_in, H)
    X = np.zeros((x.shape[0], k_in))
    X[:, : x.shape[1]] = x
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    Z3 = H2 @ W3 + b3
    H3 = gelu(Z3)
    y = H3 @ w4 + b4[0]
    return y, (X, Z1, H1, Z2, H2, Z3, H3)


def loss_and_grad(params, x, t, k_in, H):
    W1, b1, W2, b2, W3, b3, w4, b4 = split(params, k_in, H)
    y, (X, Z1, H1, Z2, H2, Z3, H3) = forward(params, x, k_in, H)
    B = x.shape[0]
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / B
    g_w4 = H3.T @ dy
    g_b4 = np.array([dy.sum()])
    dZ3 = np.outer(dy, w4) * gelu_grad(Z3)
    g_W3 = H2.T @ dZ3
    g_b3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3.T) * gelu_grad(Z2)
    g_W2 = H1.T @ dZ2
    g_b2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * gelu_grad(Z1)
    g_W1 = X.T @ dZ1
    g_b1 = dZ1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_W3.ravel(), g_b3, g_w4, g_b4])
    return loss, g
