"""The selector network's HIP kernels vs the FP64 oracle (runs on an MI355X).

Tolerances are FP32-vs-FP64: forward probabilities to 2e-6 absolute, gradients
to 1e-4 normwise relative, one Adam step to 1e-6 absolute."""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
from eigen_amd.selector import Selector
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _rand(seed, B):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 0.4, M.N_PARAMS).astype(np.float32)
    X = rng.uniform(0, 1, (B, 7)).astype(np.float32)
    Q = rng.uniform(0, 1, (B, 24))
    Q = (Q / Q.sum(axis=1, keepdims=True)).astype(np.float32)
    return P, X, Q


def test_features_match_oracle(gpu):
    for shape in [(0, 0, 0, 64, 64, 64), (1, 1, 0, 16384, 77, 1000), (0, 0, 1, 5000, 8192, 300)]:
        d, ta, tb, m, n, k = shape
        x = np.empty(7, np.float32)
        E.check(E.lib().eh_mlp_features(d, ta, tb, m, n, k, x.ctypes.data_as(C.c_void_p)))
        assert np.allclose(x, M.features(d, ta, tb, m, n, k), atol=1e-7)


def test_forward_matches_oracle(gpu):
    import torch
    P, X, _ = _rand(1, 53)
    s = Selector(capacity=64)
    s.set_params(P)
    probs = torch.empty((53, 24), dtype=torch.float32, device="cuda")
    dX = _dev(X)  # kept alive until the kernels have run
    E.check(E.lib().eh_mlp_forward(s.h, dX.data_ptr(), 53, probs.data_ptr(), None))
    torch.cuda.synchronize()
    want = M.forward(P.astype(np.float64), X.astype(np.float64))[2]
    assert np.max(np.abs(probs.cpu().numpy() - want)) < 2e-6
    # argmax prediction agrees wherever the top two are not tied within FP32
    pred = s.predict(X)
    top2 = np.sort(want, axis=1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) > 1e-5
    assert np.array_equal(pred[clear], np.argmax(want, axis=1)[clear])
    s.close()


def test_backward_matches_oracle_and_finite_differences(gpu):
    import torch
    P, X, Q = _rand(2, 41)
    s = Selector(capacity=64)
    s.set_params(P)
    g = torch.empty(M.N_PARAMS, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    dX, dQ = _dev(X), _dev(Q)  # kept alive until the kernels have run
    E.check(E.lib().eh_mlp_backward(s.h, dX.data_ptr(), dQ.data_ptr(), 41, g.data_ptr(), loss.data_ptr(), None))
    torch.cuda.synchronize()
    P64, X64, Q64 = P.astype(np.float64), X.astype(np.float64), Q.astype(np.float64)
    want = M.backward(P64, X64, Q64)
    got = g.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - want) <= 1e-4 * np.linalg.norm(want)
    assert abs(float(loss.cpu()[0]) - M.loss(P64, X64, Q64)) < 1e-5 * max(1.0, M.loss(P64, X64, Q64))
    idx = [0, 100, 447, M.OFF_B1 + 10, M.OFF_W2 + 500, M.OFF_B2 + 7]
    fd = M.finite_difference(P64, X64, Q64, idx)
    assert np.allclose(got[idx], fd, rtol=1e-3, atol=1e-6)
    s.close()


def test_adam_step_matches_oracle(gpu):
    import torch
    P, X, Q = _rand(3, 16)
    s = Selector(capacity=16)
    s.set_params(P)
    G = np.random.default_rng(4).normal(0, 1e-2, M.N_PARAMS).astype(np.float32)
    dG = _dev(G)
    E.check(E.lib().eh_mlp_adam_step(s.h, dG.data_ptr(), 0.01, 0.9, 0.999, 1e-8, 1e-3, None))
    torch.cuda.synchronize()
    want, _, _ = M.adam_step(P.astype(np.float64), G.astype(np.float64), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS),
                             1, lr=0.01, wd=1e-3)
    assert np.max(np.abs(s.params() - want)) < 1e-6
    s.close()


def test_training_fits_separable_targets(gpu):
    # targets: the configuration index is a function of the features
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (512, 7)).astype(np.float32)
    lab = (np.floor(X[:, 0] * 4).astype(int) * 6 + np.floor(X[:, 2] * 6).astype(int)) % 24
    Q = np.full((512, 24), 0.001, np.float32)
    Q[np.arange(512), lab] = 1.0
    Q /= Q.sum(axis=1, keepdims=True)
    s = Selector(capacity=512)
    curve = s.train(X, Q, epochs=1500, lr=2e-2, wd=0.0)
    assert curve[-1] < 0.5 * curve[0]
    assert np.mean(s.predict(X) == lab) > 0.8
    s.close()


def test_planner_uses_installed_network(gpu):
    P, _, _ = _rand(6, 1)
    s = Selector(capacity=8)
    s.set_params(P)
    s.install()
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_MLP)
    try:
        shapes = [(0, 0, 0, 4096, 4096, 4096), (1, 1, 1, 300, 9000, 70), (0, 1, 0, 64, 64, 16384)]
        X = np.array([M.features(*sh) for sh in shapes], np.float32)
        dev_pred = s.predict(X)
        host_pred = [E.plan(d, ta, tb, m, n, k) for d, ta, tb, m, n, k in shapes]
        assert list(dev_pred) == host_pred
    finally:
        E.set_tune_mode(prev)
        E.check(E.lib().eh_planner_set_mlp(None, None))
        s.close()
