"""CPU restatement of the learned substamp screen's network — test
infrastructure only (tests/ and the bench's CPU leg import it; the library
never does).

The reference has no learned component: its clip-refit loop
(src/cdkscUtil.cpp:585-684) is the method, and the network here only decides
which substamps to replace before that loop judges them.  What this file
restates is the library's own network as the header include/xbach.h defines
it (XB_MLP_*): inputs x = (features - mean) / sd, h = relu(W1 x + b1),
z = w2 . h + b2, p = 1 / (1 + exp(-z)), W1 and w2 rounded to FP16;
loss L = sum_k wt_k * BCE(p_k, y_k) / n.  Everything is computed in float64
so it checks the GPU's FP32 kernels, and its gradient is pinned by central
finite differences (tests/test_mlp.py).
"""
from __future__ import annotations

import numpy as np

NFEAT, NHID = 32, 64
MU, SD, W1 = 0, 32, 64
B1 = W1 + NHID * NFEAT
W2 = B1 + NHID
B2 = W2 + NHID
TAU = B2 + 1
NPARAM = TAU + 1
NGRAD = NHID * NFEAT + 2 * NHID + 1


def unpack(model, fp16: bool = True):
    m = np.asarray(model, np.float64)
    w1 = m[W1:B1].reshape(NHID, NFEAT)
    w2 = m[W2:B2]
    if fp16:
        w1 = w1.astype(np.float16).astype(np.float64)
        w2 = w2.astype(np.float16).astype(np.float64)
    return m[MU:SD], m[SD:W1], w1, m[B1:W2], w2, m[B2]


def forward(model, X, fp16: bool = True):
    """(p, z, pre-activations, normalised inputs) for features X[n][32]."""
    mu, sd, w1, b1, w2, b2 = unpack(model, fp16)
    x = (np.asarray(X, np.float64) - mu) / sd
    a = x @ w1.T + b1
    h = np.maximum(a, 0.0)
    z = h @ w2 + b2
    return 1.0 / (1.0 + np.exp(-z)), z, a, x


def loss(model, X, y, wt, fp16: bool = True) -> float:
    _, z, _, _ = forward(model, X, fp16)
    y = np.asarray(y, np.float64)
    per = np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))
    return float(np.sum(np.asarray(wt, np.float64) * per) / len(z))


def grad(model, X, y, wt, fp16: bool = True) -> np.ndarray:
    """dL/d(W1, b1, w2, b2) in that order (the trainable part of the model)."""
    p, z, a, x = forward(model, X, fp16)
    _, _, _, _, w2, _ = unpack(model, fp16)
    n = len(z)
    dz = np.asarray(wt, np.float64) * (p - np.asarray(y, np.float64)) / n
    h = np.maximum(a, 0.0)
    dpre = np.outer(dz, w2) * (a > 0.0)
    return np.concatenate([(dpre.T @ x).ravel(), dpre.sum(axis=0), h.T @ dz, [dz.sum()]])
