"""CPU restatement of the cost-model network (test infrastructure only: used
by tests/ as the checker of the device kernels, never by the product).

The network is the one litequantumcomputingsimulator_amd/csrc/device/mlp.hip
trains: F -> 64 -> 64 -> 1 with ReLU, FP64, parameters packed as W1 [F][64]
(input-major), b1 [64], W2 [64][64], b2 [64], w3 [64], b3; loss = mean
squared error. It has no counterpart in the reference.

Pinned by central finite differences of loss() (tests/test_oracle.py),
which fixes every term of grad() independently of how it is written here.
"""
from __future__ import annotations

import numpy as np

H = 64


def unpack(p: np.ndarray, F: int):
    k = 0
    W1 = p[k:k + F * H].reshape(F, H); k += F * H
    b1 = p[k:k + H]; k += H
    W2 = p[k:k + H * H].reshape(H, H); k += H * H
    b2 = p[k:k + H]; k += H
    w3 = p[k:k + H]; k += H
    b3 = p[k]; k += 1
    assert k == len(p)
    return W1, b1, W2, b2, w3, b3


def num_params(F: int) -> int:
    return F * H + H + H * H + H + H + 1


def forward(p: np.ndarray, X: np.ndarray) -> np.ndarray:
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    A1 = np.maximum(X @ W1 + b1, 0.0)
    A2 = np.maximum(A1 @ W2 + b2, 0.0)
    return A2 @ w3 + b3


def loss(p: np.ndarray, X: np.ndarray, t: np.ndarray) -> float:
    d = forward(p, X) - t
    return float(np.mean(d * d))


def grad(p: np.ndarray, X: np.ndarray, t: np.ndarray) -> np.ndarray:
    F = X.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(p, F)
    Z1 = X @ W1 + b1
    A1 = np.maximum(Z1, 0.0)
    Z2 = A1 @ W2 + b2
    A2 = np.maximum(Z2, 0.0)
    y = A2 @ w3 + b3
    dy = 2.0 * (y - t) / len(t)
    dw3 = A2.T @ dy
    db3 = dy.sum()
    dZ2 = np.outer(dy, w3) * (Z2 > 0)
    dW2 = A1.T @ dZ2
    db2 = dZ2.sum(axis=0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    dW1 = X.T @ dZ1
    db1 = dZ1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
