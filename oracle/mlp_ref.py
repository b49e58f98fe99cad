"""CPU restatement of the decode-cost MLP (TEST INFRASTRUCTURE ONLY).

float64 numpy forward and backward for the parameter layout of
compression_amd/csrc/mlp.hip: W1[F][H] b1[H] W2[H][H] b2[H] w3[H] b3, ReLU
hidden layers, loss mean((y - t)^2).  The tests pin it with central finite
differences and check the HIP gradients against it.
"""
from __future__ import annotations

import numpy as np

F, H = 6, 32
P = F * H + H + H * H + H + H + 1


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ w3 + b3


def loss(p, x, t):
    r = forward(p, x) - np.asarray(t, np.float64)
    return float(np.mean(r * r))


def grad(p, x, t):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ w3 + b3
    dy = 2.0 * (y - t) / n
    g_b3 = dy.sum()
    g_w3 = h2.T @ dy
    d2 = np.outer(dy, w3) * (a2 > 0)
    g_b2 = d2.sum(0)
    g_W2 = h1.T @ d2
    d1 = (d2 @ W2.T) * (a1 > 0)
    g_b1 = d1.sum(0)
    g_W1 = x.T @ d1
    return np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])


def init(seed: int, target_mean: float = 0.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    o = 0
    p[o:o + F * H] = rng.normal(0, np.sqrt(2.0 / F), F * H); o += F * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2.0 / H), H * H); o += H * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H] = rng.normal(0, np.sqrt(1.0 / H), H); o += H
    p[o] = target_mean
    return p
