"""GPU: the closure network's BF16-MFMA forward and backward kernels against
the numpy oracle with the same rounding points (oracle/mlp_oracle.py), and
its training (Adam) on a regression with a known answer."""
import numpy as np
import pytest

import arkamd
import mlp_oracle as M

pytestmark = pytest.mark.gpu


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 10)).astype(np.float32) * 2.0 + 0.5
    T = (np.tanh(0.35 * (X[:, :6] - 0.5)) + 0.15 * X[:, 6:7] * X[:, 7:8]).astype(np.float32)
    return X, T


NORM = (np.full(10, 0.5), np.full(10, 2.0), np.zeros(6), np.full(6, 0.5))


def make(seed=3):
    m = arkamd.MLP(seed=seed)
    m.set_normalization(*NORM)
    return m


@pytest.mark.parametrize("n", [1, 31, 64, 1000])
def test_forward_vs_oracle(n):
    X, _ = data(n)
    with make() as m:
        p = m.params().astype(np.float64)
        y = m.predict(X)
        want = M.predict(p, X.astype(np.float64), NORM, bf=True)
        scale = np.abs(want).max()
        assert np.abs(y - want).max() < 2e-3 * scale
        assert np.abs(y - want).mean() < 2e-5 * scale


@pytest.mark.parametrize("n", [32, 777, 4096])
def test_gradient_vs_oracle(n):
    X, T = data(n, seed=n)
    with make() as m:
        p = m.params().astype(np.float64)
        g, loss = m.gradient(X, T)
        l0, g0, _ = M.forward_backward(p, X.astype(np.float64), T.astype(np.float64), NORM, bf=True)
        assert abs(loss - l0) < 1e-4 * l0
        err = np.linalg.norm(g - g0) / np.linalg.norm(g0)
        assert err < 2e-3, err
        # and against the textbook float64 gradient (within the BF16 rounding)
        _, g64, _ = M.forward_backward(p, X.astype(np.float64), T.astype(np.float64), NORM, bf=False)
        assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2


def test_fresh_networks_match_the_oracle():
    """Networks created back to back each hold exactly their initial weights
    (the BF16 copies' zero padding is written on the network's own stream,
    before the copies themselves)."""
    X, T = data(33, seed=5)
    for seed in range(8):
        with make(seed=seed) as m:
            p = m.params().astype(np.float64)
            _, loss = m.gradient(X, T)
            l0 = M.forward_backward(p, X.astype(np.float64), T.astype(np.float64), NORM, bf=True)[0]
            assert abs(loss - l0) < 1e-5 * l0, (seed, loss, l0)


def test_training_converges_and_is_deterministic():
    X, T = data(20000, seed=9)
    runs = []
    for _ in range(2):
        with make(seed=7) as m:
            m.set_data(X, T)
            p0 = m.params().astype(np.float64)
            l0 = M.forward_backward(p0, X.astype(np.float64), T.astype(np.float64), NORM, bf=True)[0]
            m.train(600, batch=4096, lr=2e-3)
            p = m.params()
            runs.append(p)
            _, l1 = m.gradient(X, T)
    assert np.array_equal(runs[0], runs[1])
    assert l1 < 0.05 * l0, (l0, l1)


def test_one_adam_step_vs_oracle():
    """One training step on the whole data set (batch drawn with replacement,
    so compare through the gradient the step used: set_data with the same
    rows and a batch size of n, then check the update direction)."""
    X, T = data(256, seed=4)
    with make() as m:
        p0 = m.params().astype(np.float64)
        g, _ = m.gradient(X, T)
        m.set_data(X, T)
        m.train(1, batch=256, lr=1e-3)
        p1 = m.params().astype(np.float64)
    step = p1 - p0
    # Adam's first step moves every parameter by lr * sign(g) (|g| >> eps)
    big = np.abs(g) > 1e-5
    assert np.all(np.abs(np.abs(step[big]) - 1e-3) < 2e-4)
