"""CPU oracle of the learned position model -- TEST INFRASTRUCTURE ONLY.

The learned index has no counterpart in the reference.  This is the model's
mathematics in float64 numpy, used to check the GPU training kernels' gradients and the
leaf routing; it is itself pinned by central finite differences (tests/test_oracle.py).

Parameter layout (321 floats): W1[16], b1[16], W2[16][16] (row k = hidden-2 unit k),
b2[16], w3[16], b3.  Model: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = w3 . h2 + b3.
Loss over a batch of B samples (x_i, t_i): mean (y_i - t_i)^2.
"""
from __future__ import annotations

import numpy as np

H = 16
NPARAM = H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    W1 = p[0:H]
    b1 = p[H:2 * H]
    W2 = p[2 * H:2 * H + H * H].reshape(H, H)
    b2 = p[2 * H + H * H:3 * H + H * H]
    w3 = p[3 * H + H * H:4 * H + H * H]
    b3 = p[4 * H + H * H]
    return W1, b1, W2, b2, w3, b3


def identity_params():
    """The model's initialisation: h1_j = relu(x - j/16), W2 = I, y = h2_0 (= x on [0, 1])."""
    p = np.zeros(NPARAM, np.float64)
    p[0:H] = 1.0
    p[H:2 * H] = -np.arange(H) / H
    p[2 * H:2 * H + H * H] = np.eye(H).ravel()
    p[3 * H + H * H] = 1.0
    return p


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    z1 = np.outer(x, W1) + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ w3 + b3


def loss(p, x, t):
    r = forward(p, x) - np.asarray(t, np.float64)
    return float(np.mean(r * r))


def gradient(p, x, t):
    """Analytic gradient of loss(p, x, t) with respect to p (float64), plus the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    B = x.size
    z1 = np.outer(x, W1) + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    r = y - t
    dy = 2.0 * r / B
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    d2 = np.where(z2 > 0, np.outer(dy, w3), 0.0)
    g_W2 = d2.T @ h1
    g_b2 = d2.sum(0)
    dh1 = d2 @ W2
    d1 = np.where(z1 > 0, dh1, 0.0)
    g_W1 = d1.T @ x
    g_b1 = d1.sum(0)
    g = np.concatenate([g_W1, g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return g, float(np.mean(r * r))


def leaf_of(y, L):
    f = np.asarray(y, np.float64) * L
    return np.clip(np.floor(np.nan_to_num(f, nan=0.0)), 0, L - 1).astype(np.int64)
