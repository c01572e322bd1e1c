"""Pins the FP64 restatement of the selector network (CPU only): finite
differences fix the gradient, softmax identities and the all-zero network fix
the forward pass, and a hand-computed Adam step fixes the update."""
import math

import numpy as np

from oracle import mlp as M


def _rand(seed, B=9):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 0.5, M.N_PARAMS)
    X = rng.uniform(0, 1, (B, M.N_IN))
    Q = rng.uniform(0, 1, (B, M.N_OUT))
    Q /= Q.sum(axis=1, keepdims=True)
    return P, X, Q


def test_param_count():
    assert M.N_PARAMS == 2072


def test_zero_network_is_uniform():
    P = np.zeros(M.N_PARAMS)
    X = np.random.default_rng(0).uniform(size=(5, 7))
    Q = np.zeros((5, 24))
    Q[np.arange(5), [0, 3, 7, 11, 23]] = 1.0
    _, Z, Pr = M.forward(P, X)
    assert np.allclose(Pr, 1.0 / 24) and np.all(Z == 0)
    assert math.isclose(M.loss(P, X, Q), math.log(24.0), rel_tol=1e-15)
    g = M.backward(P, X, Q)
    # dz = (1/24 - q)/B, hidden activations are zero -> only b2 has a gradient
    assert np.allclose(g[M.OFF_B2:], (1.0 / 24 * 5 - Q.sum(axis=0)) / 5)
    assert not np.any(g[:M.OFF_B2])


def test_softmax_identities():
    P, X, _ = _rand(1)
    _, Z, Pr = M.forward(P, X)
    assert np.allclose(Pr.sum(axis=1), 1.0)
    P2 = P.copy()
    P2[M.OFF_B2:] += 3.7  # shifting every logit leaves p unchanged
    assert np.allclose(M.forward(P2, X)[2], Pr)


def test_gradient_matches_finite_differences():
    P, X, Q = _rand(2)
    g = M.backward(P, X, Q)
    idx = list(range(0, M.N_PARAMS, 37)) + [M.OFF_B1 + 3, M.OFF_B2 + 5, M.N_PARAMS - 1]
    fd = M.finite_difference(P, X, Q, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-9)
    # a planted sign error in the hidden layer would be caught
    bad = g.copy()
    bad[M.OFF_W1:M.OFF_B1] *= -1
    assert not np.allclose(bad[idx], fd, rtol=1e-5, atol=1e-9)


def test_adam_first_step_closed_form():
    # at step 1 the bias-corrected update is lr * g / (|g| + eps) = lr * sign(g)
    P = np.ones(M.N_PARAMS)
    G = np.linspace(-1, 1, M.N_PARAMS) + 1e-3
    P1, m, v = M.adam_step(P, G, np.zeros_like(G), np.zeros_like(G), 1, lr=0.01)
    assert np.allclose(P1, 1.0 - 0.01 * np.sign(G), atol=1e-7)
    # decoupled weight decay touches weights only
    P2, _, _ = M.adam_step(P, np.zeros_like(G), np.zeros_like(G), np.zeros_like(G), 1, lr=0.1, wd=0.5)
    assert np.allclose(P2[M.OFF_W1:M.OFF_B1], 0.95) and np.allclose(P2[M.OFF_B1:M.OFF_W2], 1.0)
    assert np.allclose(P2[M.OFF_W2:M.OFF_B2], 0.95) and np.allclose(P2[M.OFF_B2:], 1.0)


def test_features():
    x = M.features(1, True, False, 1024, 64, 16384)
    assert np.allclose(x, [10 / 14, 6 / 14, 1.0, 0, 1, 1, 0])
