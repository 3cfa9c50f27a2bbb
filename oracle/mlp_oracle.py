"""Test infrastructure only: numpy restatement of the closure network
(ark_mlp: 10 -> 64 -> 64 -> 6, SiLU, mean squared error) with its forward
pass, its backward pass and Adam.  Two variants:

* ``bf16=False``: everything in float64, the textbook network.  Its
  gradient is pinned by central finite differences of the loss
  (tests/test_mlp_oracle.py).
* ``bf16=True``: the rounding points of the HIP kernels -- weights, inputs,
  hidden activations and the backward products' operands rounded to
  bfloat16 (round to nearest even) -- with float64 sums in place of the
  matrix cores' float32 sums.  The GPU's outputs and gradients are checked
  against this variant.

Parameter vector layout: W1[64][10] b1[64] W2[64][64] b2[64] W3[6][64] b3[6].
Nothing in the product imports this module.
"""
from __future__ import annotations

import numpy as np

NI, NH, NO = 10, 64, 6
SIZES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NO, NH)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    u = (u + (((u >> 16) & 1) + 0x7FFF)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward_backward(p, X, T, norm=None, bf=False):
    """Loss, gradient (packed) and outputs (normalised units) over all rows."""
    P = unpack(p)
    n = X.shape[0]
    im, istd, om, ostd = norm if norm is not None else (np.zeros(NI), np.ones(NI), np.zeros(NO), np.ones(NO))
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    if bf:
        xn = ((np.float32(1) * X.astype(np.float32) - np.float32(im)) * (np.float32(1) / np.float32(istd)))
        tn = ((T.astype(np.float32) - np.float32(om)) * (np.float32(1) / np.float32(ostd))).astype(np.float64)
    else:
        xn = (X - im) / istd
        tn = (T - om) / ostd
    x = r(xn)                                          # n x 10
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    z1 = x @ W1.T + P["b1"]
    h1 = r(z1 * sig(z1))
    z2 = h1 @ W2.T + P["b2"]
    h2 = r(z2 * sig(z2))
    y = h2 @ W3.T + P["b3"]
    e = y - tn
    loss = (e * e).sum() / (NO * n)
    dy = 2.0 * e / (NO * n)
    dyb = r(dy)
    dz2 = (dyb @ W3) * (sig(z2) * (1 + z2 * (1 - sig(z2))))
    dz2b = r(dz2)
    dz1 = (dz2b @ W2) * (sig(z1) * (1 + z1 * (1 - sig(z1))))
    dz1b = r(dz1)
    g = {"W3": dyb.T @ h2, "b3": dy.sum(0), "W2": dz2b.T @ h1, "b2": dz2.sum(0),
         "W1": dz1b.T @ x, "b1": dz1.sum(0)}
    return loss, pack(g), y


def predict(p, X, norm=None, bf=True):
    _, _, y = forward_backward(p, X, np.zeros((X.shape[0], NO)), norm, bf)
    im, istd, om, ostd = norm if norm is not None else (np.zeros(NI), np.ones(NI), np.zeros(NO), np.ones(NO))
    return y * ostd + om


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
