"""CPU: the closure-network oracle (oracle/mlp_oracle.py).  Its float64
gradient is pinned by central finite differences of its loss; its bfloat16
rounding by exact cases."""
import numpy as np

import mlp_oracle as M


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 10))
    T = np.tanh(X[:, :6] * 0.7) + 0.3 * X[:, 6:7] * X[:, 7:8]
    return X, T


def params(seed=1):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(size=(64, 10)) / np.sqrt(10), "b1": 0.1 * rng.normal(size=64),
         "W2": rng.normal(size=(64, 64)) / 8.0, "b2": 0.1 * rng.normal(size=64),
         "W3": rng.normal(size=(6, 64)) / 8.0, "b3": 0.1 * rng.normal(size=6)}
    return M.pack(d)


def test_gradient_matches_finite_differences():
    X, T = data(40)
    p = params()
    loss, g, _ = M.forward_backward(p, X, T)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(M.NPARAM, 40, replace=False), [0, 639, 640, 4800, 4863, 5253]])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (M.forward_backward(pp, X, T)[0] - M.forward_backward(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 + 1e-5 * abs(g[q]), (q, fd, g[q])


def test_bf16_rounding():
    # bfloat16 keeps 7 fraction bits: spacing 2^-7 at 1.0
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-5],
                 dtype=np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 and 1 + 3*2^-8 are ties: to the even neighbour (1.0 and 1 + 2^-6)
    assert list(r[:5]) == [1.0, 1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6]
    assert r[5] == -2.5
    assert abs(r[6] - 3.0e-5) / 3.0e-5 < 2 ** -8


def test_adam_first_step_is_lr_sign():
    p, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    g = np.array([0.5, -2.0, 1e-3])
    p1, _, _ = M.adam(p, m, v, g, 1, 1e-3)
    assert np.allclose(p1, -1e-3 * np.sign(g), rtol=1e-4)


def test_training_reduces_loss():
    X, T = data(512)
    p = params()
    m = v = np.zeros_like(p)
    l0 = M.forward_backward(p, X, T)[0]
    for t in range(1, 201):
        _, g, _ = M.forward_backward(p, X, T)
        p, m, v = M.adam(p, m, v, g, t, 3e-3)
    assert M.forward_backward(p, X, T)[0] < 0.3 * l0
