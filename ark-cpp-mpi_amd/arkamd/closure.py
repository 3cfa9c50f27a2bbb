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
    """A network trained on inputs X and targets T; "norm" holds the input and
    output normalisation (mean, std of X, then of T) it was trained with."""
    m = MLP(seed=seed)
    norm = (X.mean(0), X.std(0) + 1e-12, T.mean(0), T.std(0) + 1e-12)
    m.set_normalization(*norm)
    m.set_data(X, T)
    _, l0 = m.gradient(X[: min(len(X), 65536)], T[: min(len(T), 65536)])
    t0 = time.perf_counter()
    last = m.train(steps, batch=batch, lr=lr)
    wall = time.perf_counter() - t0
    _, l1 = m.gradient(X[: min(len(X), 65536)], T[: min(len(T), 65536)])
    return {"mlp": m, "norm": norm, "loss0": l0, "loss_end": l1, "last_batch_loss": last, "wall": wall}


def set_amplitude(tr: Dict, alpha: float) -> None:
    """Scale the trained network's stresses by alpha (through its output
    normalisation: the stress is alpha (y std_T + mean_T) for network output y)."""
    xm, xs, tm, ts = tr["norm"]
    tr["mlp"].set_normalization(xm, xs, alpha * tm, alpha * ts)
    tr["alpha"] = alpha


def calibrate(d: Dict, tr: Dict, vis: float, eps: float, seed: int, delta: float, tau: float = 0.01,
              alphas=(1.0, 1.25, 1.5, 1.75, 2.0, 2.5, 3.0)) -> Dict:
    """A-posteriori amplitude of the closure, chosen on the training data.
    A network fitted to the lag by least squares predicts its conditional mean
    given the local gradients, which explain only part of it; the coarse run
    then needs a stronger correction than that mean.  For each candidate
    alpha, one coarse run from the filtered initial field passes through the
    start times of the correction windows (d["windows"]), and its spectrum is
    compared there with the filtered fine field it was trained on: the mean
    error over shells 1..N/4 and over the windows that start in the last
    quarter of the fine run (at least the last window).  Earlier windows see
    a cascade still filling the high shells (at N/4, 1e-7 to 4e-5 of the peak
    over the first half of the 512^3 run, 1e-4 to 1e-3 over its second half),
    whose relative errors say little about the developed flow.  The
    alpha with the least error among the runs that stay finite is set on the
    network.  The fine run's final state, against which the closure is
    judged, takes no part."""
    w = d["windows"]
    times = [x["t_n"] for x in w]
    nc = d["ub0"].shape[-1]
    later = [q for q in range(len(w)) if w[q]["t_n"] >= 0.75 * d["t_end"]] or [len(w) - 1]
    table = []
    for a in alphas:
        set_amplitude(tr, a)
        r = les(d["ub0"], times, vis, eps, seed, mlp=tr["mlp"], delta=delta, tau=tau)
        errs = [spectrum_error(r["E_at"][q], np.asarray(w[q]["E_ref"]), nc // 4) for q in later]
        table.append({"alpha": a, "nonfinite": r["nonfinite"], "err_windows": float(np.mean(errs))})
    ok = [t for t in table if t["nonfinite"] == 0 and np.isfinite(t["err_windows"])]
    best = min(ok, key=lambda t: t["err_windows"])["alpha"] if ok else 1.0
    set_amplitude(tr, best)
    return {"alpha": best, "candidates": table}


def les(ub0: np.ndarray, t_end, vis: float, eps: float, seed: int, mlp: Optional[MLP] = None,
        delta: float = 0.0, tau: float = 0.01, kcut: int = 0) -> Dict:
    """A coarse run from the filtered field ub0 to t_end with the same
    forcing; with mlp, the learned closure, applied on the shells >= kcut only
    when kcut > 0 (the scale-selective variant).  t_end may be a list of
    increasing times: the run then also records its spectrum on reaching each
    ("E_at"), and "wall" excludes those spectra."""
    times = list(t_end) if np.ndim(t_end) else [t_end]
    N3, N2, N1 = ub0.shape[1:]
    s = Solver(N1, N2, N3, VIS=vis, precision=PREC_MIXED)
    s.init_velocity(ub0)
    s.set_forcing(seed + 1, eps, tau)
    if mlp is not None:
        s.set_closure(mlp, delta)
        if kcut > 0:
            s.set_closure_cutoff(kcut)
    s.synchronize()
    n, wall, E_at = 0, 0.0, []
    for t in times:
        t0 = time.perf_counter()
        n += run_to(s, t)
        s.synchronize()
        wall += time.perf_counter() - t0
        if len(times) > 1:
            E_at.append(s.spectrum(N1 // 2 + 1))
    d = s.diagnostics(allow_nonfinite=True)
    out = {"steps": n, "time": d["time"], "wall": wall, "E": s.spectrum(N1 // 2 + 1), "ke": d["ke_physical"],
           "eps_resolved": s.dissipation(), "nonfinite": d["nonfinite"], "E_at": E_at}
    s.close()
    return out


def stress_from_force(F: np.ndarray, dx: float) -> np.ndarray:
    """A symmetric stress field tau (6, N3, N2, N1; 11 22 33 12 13 23) whose
    acceleration as the closure applies it, -sum_j D_j tau_ij with D_j the
    periodic central difference (tau(i+1) - tau(i-1)) / (2 dx), is the given
    acceleration F (3, N3, N2, N1), except for the modes that operator cannot
    produce (every k_j dx in {0, pi}), which are dropped.  With
    tau_ij = D_j psi_i + D_i psi_j, and D_j of symbol i s_j,
    s_j = sin(k_j dx) / dx, the condition reads (|s|^2 I + s s^T) psi = F,
    which Sherman-Morrison inverts: psi = (F - s (s.F) / (2|s|^2)) / |s|^2."""
    shp = F.shape[1:]
    ks = [np.fft.fftfreq(n, d=1.0 / n) for n in shp]      # integer wavenumbers per array axis (z, y, x)
    kz, ky, kx = np.meshgrid(*ks, indexing="ij")
    L = [n * dx for n in shp]
    s = [np.sin(2 * np.pi * kx / L[2] * dx) / dx, np.sin(2 * np.pi * ky / L[1] * dx) / dx,
         np.sin(2 * np.pi * kz / L[0] * dx) / dx]          # component order x, y, z
    s2 = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    ok = s2 > 1e-12 * s2.max()
    Fh = [np.fft.fftn(F[m]) for m in range(3)]
    sF = s[0] * Fh[0] + s[1] * Fh[1] + s[2] * Fh[2]
    inv = np.where(ok, 1.0 / np.where(ok, s2, 1.0), 0.0)
    psi = [(Fh[m] - s[m] * sF * 0.5 * inv) * inv for m in range(3)]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    return np.stack([np.real(np.fft.ifftn(1j * (s[b] * psi[a] + s[a] * psi[b]))) for a, b in pairs])


def taylor_reynolds(ke_physical: float, eps: float, vis: float, volume: float) -> float:
    """Re_lambda = u' lambda / nu with u' the per-component rms velocity
    (u'^2 = 2 KE / (3 V)) and the Taylor microscale lambda = sqrt(15 nu u'^2 / eps)."""
    up2 = 2.0 * ke_physical / (3.0 * volume)
    return float(up2 * np.sqrt(15.0 / (vis * eps)))


def correction_pairs(N: int, steps: int, pairs: int, width: int, vis: float, eps: float, seed: int,
                     k0: float = 4.0, urms: float = 1.0, tau: float = 0.01, horizon: int = 8,
                     scheme: Optional[int] = None, series_every: int = 0) -> Dict:
    """Training pairs for a closure of the coarse solver itself (a posteriori
    targets), from one forced fine run.  `pairs` times the fine state is
    box-filtered (width w) to the coarse grid, velocity and density, and a
    coarse context started from that filtered state at the same TIME takes
    `horizon` steps with the same forcing.  The fine run continues to the
    coarse run's final time (the filtered velocities of the two fine steps
    around it, interpolated in time).  The acceleration the coarse run lacked,
        F = (ubar_fine(t_c) - u_coarse(t_c)) / (t_c - t_n),
    holds its subgrid stress and its own discretisation error together; the
    stress whose closure force is F (stress_from_force) is the target "T",
    the filtered velocity gradient and the filter width at t_n the inputs "X".
    "T_sgs" holds the exact subgrid stresses of the same states (the a-priori
    targets).  "wall" is the fine run's own stepping time; "series" its
    kinetic energy, dissipation and Re_lambda every `series_every` steps."""
    from . import SCHEME_REFERENCE
    scheme = SCHEME_REFERENCE if scheme is None else scheme
    Nc = N // width
    s = Solver(N, N, N, VIS=vis, precision=PREC_MIXED, scheme=scheme)
    s.init_turbulence(seed, k0, urms)
    s.set_forcing(seed + 1, eps, tau)
    c = Solver(Nc, Nc, Nc, VIS=vis, precision=PREC_MIXED, scheme=scheme)
    c.set_forcing(seed + 1, eps, tau)
    volume = s.dx[0] * s.dx[1] * s.dx[2] * N ** 3
    dxc = width * s.dx[0]
    ub0 = s.filter_box(width)[0]
    at = [int(round(steps * (q + 1) / (pairs + 1))) for q in range(pairs)]
    Xs, Ts, Ss, info, series = [], [], [], [], []
    state = {"done": 0, "wall": 0.0}

    def fine(k):
        s.synchronize()
        t0 = time.perf_counter()
        s.step(k)
        s.synchronize()
        state["wall"] += time.perf_counter() - t0
        for q in range(state["done"] + 1, state["done"] + k + 1):
            if series_every and q % series_every == 0 and q == state["done"] + k:
                d = s.diagnostics(allow_nonfinite=True)
                e = s.dissipation()
                series.append({"step": q, "time": d["time"], "ke": d["ke_physical"], "eps": e,
                               "Re_lambda": taylor_reynolds(d["ke_physical"], e, vis, volume)})
        state["done"] += k

    def fine_to(target):
        while state["done"] < target:
            nxt = target
            if series_every:
                nxt = min(target, (state["done"] // series_every + 1) * series_every)
            fine(nxt - state["done"])

    for target in at:
        fine_to(target)
        ub, tsgs, grad = s.filter_box(width)
        rho = s.filter_density(width)
        t_n = s.clock()["time"]
        c.init_state(ub, rho)
        c.restart({"time": t_n, "nstep": 0, "dt": 0.0, "dt_next": 0.0, "dt_next_known": 0})
        c.step(horizon)
        t_c = c.clock()["time"]
        uc = np.stack([c.download(n)[1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")])
        while True:                          # fine steps until the next one passes t_c
            k = s.clock()
            if k["time"] + k["dt_next"] >= t_c:
                break
            fine(1)
        ta, ua = s.clock()["time"], s.filter_box(width)[0]
        fine(1)
        tb, ubb = s.clock()["time"], s.filter_box(width)[0]
        a = (t_c - ta) / (tb - ta)
        F = ((1.0 - a) * ua + a * ubb - uc) / (t_c - t_n)
        T = stress_from_force(F, dxc)
        n = ub[0].size
        X = np.empty((n, 10), dtype=np.float32)
        X[:, :9] = grad.reshape(9, n).T
        X[:, 9] = dxc
        Xs.append(X)
        Ts.append(T.reshape(6, n).T.astype(np.float32))
        Ss.append(tsgs.reshape(6, n).T.astype(np.float32))
        info.append({"step": state["done"], "t_n": t_n, "t_c": t_c, "F_rms": float(np.sqrt((F ** 2).mean())),
                     "E_ref": spectrum_of(ub, vis)})
    fine_to(steps)
    d = s.diagnostics(allow_nonfinite=True)
    e_end = s.dissipation()
    out = {"X": np.concatenate(Xs), "T": np.concatenate(Ts), "T_sgs": np.concatenate(Ss), "windows": info,
           "t_end": d["time"], "ub0": ub0, "ub_end": s.filter_box(width)[0], "dx": s.dx[0], "eps_end": e_end,
           "ke_end": d["ke_physical"], "Re_lambda_end": taylor_reynolds(d["ke_physical"], e_end, vis, volume),
           "nonfinite": d["nonfinite"], "wall": state["wall"], "steps": state["done"], "series": series}
    c.close()
    s.close()
    return out


def spectrum_error(E: np.ndarray, Eref: np.ndarray, kmax: int) -> float:
    """Mean over shells 1..kmax of |E/Eref - 1|."""
    k = np.arange(1, kmax + 1)
    return float(np.mean(np.abs(E[k] / Eref[k] - 1.0)))
