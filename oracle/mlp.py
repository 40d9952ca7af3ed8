"""float64 numpy restatement of the cost-model network -- TEST INFRASTRUCTURE ONLY.

The network (features -> H -> H -> 1, ReLU after both hidden layers, mean squared error)
and its parameter layout W1[F x H], b1[H], W2[H x H], b2[H], W3[H x 1], b3 are those of
include/qsvh.h (qsvh_mlp_*).  The reference has no learned planner (its fusion is the
greedy fusion.rs:39-77), so this file is pinned by the mathematics instead: its analytic
gradients are checked against central finite differences (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np


def unpack(params: np.ndarray, F: int, H: int):
    p = np.asarray(params, dtype=np.float64)
    i = 0

    def take(n, shape):
        nonlocal i
        out = p[i: i + n].reshape(shape)
        i += n
        return out

    W1 = take(F * H, (F, H))
    b1 = take(H, (H,))
    W2 = take(H * H, (H, H))
    b2 = take(H, (H,))
    W3 = take(H, (H, 1))
    b3 = take(1, (1,))
    assert i == p.size
    return W1, b1, W2, b2, W3, b3


def forward(params, X, F: int, H: int):
    W1, b1, W2, b2, W3, b3 = unpack(params, F, H)
    X = np.asarray(X, dtype=np.float64)
    A1 = np.maximum(X @ W1 + b1, 0.0)
    A2 = np.maximum(A1 @ W2 + b2, 0.0)
    return (A2 @ W3 + b3)[:, 0], (X, A1, A2)


def loss_and_grad(params, X, y, F: int, H: int):
    """MSE = mean((p - y)^2) and its gradient w.r.t. the flat parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(params, F, H)
    p, (X, A1, A2) = forward(params, X, F, H)
    M = len(p)
    d = p - np.asarray(y, dtype=np.float64)
    loss = float(np.mean(d * d))
    dP = (2.0 / M) * d[:, None]
    gW3 = A2.T @ dP
    gb3 = dP.sum(axis=0)
    dA2 = (dP @ W3.T) * (A2 > 0)
    gW2 = A1.T @ dA2
    gb2 = dA2.sum(axis=0)
    dA1 = (dA2 @ W2.T) * (A1 > 0)
    gW1 = X.T @ dA1
    gb1 = dA1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
