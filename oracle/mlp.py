"""FP64 restatement of the tile-selector network -- TEST INFRASTRUCTURE ONLY.

Restates, in numpy float64, the network the product library trains with its
own HIP kernels: x (7) -> h = relu(W1 x + b1) (64) -> z = W2 h + b2 (24) ->
p = softmax(z), loss = mean over the batch of -sum_o q_o log p_o, and the
Adam update with decoupled weight decay on the weights.  Parameters are one
flat vector laid out W1[64][7], b1[64], W2[24][64], b2[24].

The network has no counterpart in the reference (it replaces the blocking
heuristic of Eigen/src/Core/products/GeneralBlockPanelKernel.h:201-405), so it
is pinned by mathematics, not by Eigen: central finite differences of `loss`
pin `backward`; softmax identities and the all-zero network (uniform p, loss
= ln 24, dz = 1/24 - q) pin `forward`.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_OUT = 7, 64, 24
OFF_W1 = 0
OFF_B1 = N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_OUT * N_HID
N_PARAMS = OFF_B2 + N_OUT


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(N_HID, N_IN)
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(N_OUT, N_HID)
    b2 = P[OFF_B2:N_PARAMS]
    return W1, b1, W2, b2


def features(dtype: int, ta: bool, tb: bool, m: int, n: int, k: int) -> np.ndarray:
    return np.array([math.log2(max(m, 1)) / 14.0, math.log2(max(n, 1)) / 14.0, math.log2(max(k, 1)) / 14.0,
                     1.0 if dtype == 0 else 0.0, 1.0 if dtype == 1 else 0.0, float(bool(ta)), float(bool(tb))])


def forward(P, X):
    W1, b1, W2, b2 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    H = np.maximum(X @ W1.T + b1, 0.0)
    Z = H @ W2.T + b2
    Zs = Z - Z.max(axis=1, keepdims=True)
    E = np.exp(Zs)
    Pr = E / E.sum(axis=1, keepdims=True)
    return H, Z, Pr


def loss(P, X, Q) -> float:
    _, _, Pr = forward(P, X)
    Q = np.asarray(Q, dtype=np.float64)
    with np.errstate(divide="ignore"):
        lp = np.where(Q > 0, np.log(np.maximum(Pr, 1e-300)), 0.0)
    return float(-(Q * lp).sum() / X.shape[0])


def backward(P, X, Q):
    """Analytic gradient of `loss` (flat, same layout as P)."""
    W1, b1, W2, b2 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    H, Z, Pr = forward(P, X)
    B = X.shape[0]
    dZ = (Pr - np.asarray(Q, dtype=np.float64)) / B
    dW2 = dZ.T @ H
    db2 = dZ.sum(axis=0)
    dH = (dZ @ W2) * (H > 0)
    dW1 = dH.T @ X
    db1 = dH.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2])


def finite_difference(P, X, Q, idx, h: float = 1e-6) -> np.ndarray:
    """Central differences of `loss` for the parameter indices `idx`."""
    P = np.asarray(P, dtype=np.float64)
    out = np.empty(len(idx))
    for j, i in enumerate(idx):
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        out[j] = (loss(Pp, X, Q) - loss(Pm, X, Q)) / (2 * h)
    return out


def adam_step(P, G, Mo, Ve, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8,
              wd: float = 0.0):
    P = np.asarray(P, dtype=np.float64).copy()
    m = b1 * Mo + (1 - b1) * G
    v = b2 * Ve + (1 - b2) * G * G
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    weight = np.zeros(N_PARAMS, dtype=bool)
    weight[OFF_W1:OFF_B1] = True
    weight[OFF_W2:OFF_B2] = True
    P -= lr * (mh / (np.sqrt(vh) + eps) + np.where(weight, wd * P, 0.0))
    return P, m, v
