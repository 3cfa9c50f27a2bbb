"""Learned-closure large-eddy simulation of forced homogeneous turbulence.

The plain fine-grid CABARET run (DNS) generates its own training data: at
snapshots, the box filter of width w gives, per coarse cell, the filtered
velocity gradient and the filter width (network inputs) and the exact
subgrid stresses (targets).  The network (arkamd.MLP) is trained on them on
the GPU.  The coarse runs start from the filtered initial field, with the
same forcing, with and without the closure, and are judged against the
filtered DNS by the kinetic-energy spectrum and the dissipation rate.
Every field operation runs in the library (libark_amd); this module only
sequences the calls and does bookkeeping on small arrays.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np

from . import MLP, PREC_MIXED, Solver


def pairs_from(s: Solver, width: int):
    """Training inputs X (n x 10: d_j ubar_i, component 3i + j, and the filter
    width) and targets T (n x 6: tau 11 22 33 12 13 23) of the filtered field,
    and the filtered velocity (3, Nc3, Nc2, Nc1)."""
    ub, tau, grad = s.filter_box(width)
    n = ub[0].size
    delta = width * s.dx[0]
    X = np.empty((n, 10), dtype=np.float32)
    X[:, :9] = grad.reshape(9, n).T
    X[:, 9] = delta
    T = tau.reshape(6, n).T.astype(np.float32)
    return X, T, ub


def spectrum_of(u: np.ndarray, vis: float) -> np.ndarray:
    """Shell spectrum of a cell velocity field (3, N3, N2, N1), computed by the
    library on a context holding it."""
    N3, N2, N1 = u.shape[1:]
    with Solver(N1, N2, N3, VIS=vis, precision=PREC_MIXED) as s:
        s.init_velocity(u)
        return s.spectrum(N1 // 2 + 1)


def run_to(s: Solver, t_end: float, chunk: int = 10, max_steps: int = 10 ** 6) -> int:
    """Steps until TIME >= t_end: chunks of `chunk` steps while they stay short
    of t_end, then single steps; returns the step count."""
    n = 0
    while n < max_steps:
        d = s.diagnostics(allow_nonfinite=True)
        if d["time"] >= t_end:
            break
        k = chunk if d["time"] + 1.5 * chunk * max(d["dt"], 1e-30) < t_end else 1
        s.step(k)
        n += k
    return n


def dns(N: int, steps: int, snapshots: int, width: int, vis: float, eps: float, seed: int,
        k0: float = 4.0, urms: float = 1.0, series_every: int = 50, tau: float = 0.01) -> Dict:
    """Forced DNS; the forcing is redrawn every `tau` in time so that the
    coarse runs see the same force history."""
    s = Solver(N, N, N, VIS=vis, precision=PREC_MIXED)
    s.init_turbulence(seed, k0, urms)
    s.set_forcing(seed + 1, eps, tau)
    _, _, ub0 = pairs_from(s, width)
    Xs, Ts, series = [], [], []
    snap_at = set(int(round(steps * (q + 1) / snapshots)) for q in range(snapshots))
    t0 = time.perf_counter()
    wall = 0.0
    for n in range(1, steps + 1):
        s.step(1)
        if n % series_every == 0 or n in snap_at:
            s.synchronize()
            wall += time.perf_counter() - t0
            d = s.diagnostics(allow_nonfinite=True)
            series.append({"step": n, "time": d["time"], "ke": d["ke_physical"], "eps": s.dissipation()})
            if n in snap_at:
                X, T, ub = pairs_from(s, width)
                Xs.append(X)
                Ts.append(T)
            t0 = time.perf_counter()
    s.synchronize()
    wall += time.perf_counter() - t0
    d = s.diagnostics(allow_nonfinite=True)
    out = {"X": np.concatenate(Xs), "T": np.concatenate(Ts), "ub0": ub0, "ub_end": ub, "t_end": d["time"],
           "series": series, "wall": wall, "steps": steps, "nonfinite": d["nonfinite"],
           "eps_end": s.dissipation(), "dx": s.dx[0]}
    s.close()
    return out


def train(X: np.ndarray, T: np.ndarray, steps: int, batch: int, lr: float, seed: int = 0) -> Dict:
    m = MLP(seed=seed)
    m.set_normalization(X.mean(0), X.std(0) + 1e-12, T.mean(0), T.std(0) + 1e-12)
    m.set_data(X, T)
    _, l0 = m.gradient(X[: min(len(X), 65536)], T[: min(len(T), 65536)])
    t0 = time.perf_counter()
    last = m.train(steps, batch=batch, lr=lr)
    wall = time.perf_counter() - t0
    _, l1 = m.gradient(X[: min(len(X), 65536)], T[: min(len(T), 65536)])
    return {"mlp": m, "loss0": l0, "loss_end": l1, "last_batch_loss": last, "wall": wall}


def les(ub0: np.ndarray, t_end: float, vis: float, eps: float, seed: int, mlp: Optional[MLP] = None,
        delta: float = 0.0, tau: float = 0.01, kcut: int = 0) -> Dict:
    """A coarse run from the filtered field ub0 to t_end with the same
    forcing; with mlp, the learned closure, applied on the shells >= kcut only
    when kcut > 0 (the scale-selective variant)."""
    N3, N2, N1 = ub0.shape[1:]
    s = Solver(N1, N2, N3, VIS=vis, precision=PREC_MIXED)
    s.init_velocity(ub0)
    s.set_forcing(seed + 1, eps, tau)
    if mlp is not None:
        s.set_closure(mlp, delta)
        if kcut > 0:
            s.set_closure_cutoff(kcut)
    s.synchronize()
    t0 = time.perf_counter()
    n = run_to(s, t_end)
    s.synchronize()
    wall = time.perf_counter() - t0
    d = s.diagnostics(allow_nonfinite=True)
    out = {"steps": n, "time": d["time"], "wall": wall, "E": s.spectrum(N1 // 2 + 1), "ke": d["ke_physical"],
           "eps_resolved": s.dissipation(), "nonfinite": d["nonfinite"]}
    s.close()
    return out


def spectrum_error(E: np.ndarray, Eref: np.ndarray, kmax: int) -> float:
    """Mean over shells 1..kmax of |E/Eref - 1|."""
    k = np.arange(1, kmax + 1)
    return float(np.mean(np.abs(E[k] / Eref[k] - 1.0)))
