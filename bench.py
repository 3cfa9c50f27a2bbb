#!/usr/bin/env python3
"""Benchmark of the MI355X CABARET/ARK step: Mcell-steps/s at 256^3.

Workload: the Taylor-Green vortex on a periodic box of side 2 pi, N=256 cells
per side, c=10 (Ma=0.1), Re=1600, CFL 0.3, 500 steps, inputs resident in HBM.
One "step" is one iteration of the reference's driver loop (source/main.cpp:
82-91) over the whole box.  The headline value is the mixed-precision step;
the FP64 step (bitwise equal to the reference) is timed beside it.

Printed: ONE JSON line.  Beyond the common fields it carries
  fp64          the FP64 step's throughput and roofline
  roofline      the mixed step's fused kernel: algorithmic HBM bytes per launch
                (200 B per cell: 5 FP64 cell fields and 15 FP32 face fields,
                each read once and written once) over its average launch time,
                timed with HIP events on the library's stream
  accuracy      mixed vs FP64 after the same steps, per field (corrected
                scheme: the reference scheme goes unstable on this vortex)
  pcie_inclusive  the mixed run with its 20 state arrays downloaded and
                uploaded through the host boundary (pageable FP64 arrays),
                the transfers timed apart from the steps
  batched       64 independent boxes of uneven size and Reynolds number in one
                launch per step (graph-replayed, tile shape autotuned), half
                Taylor-Green, half forced turbulence, filtered into training pairs
  forced_turbulence_closure
                512^3 forced turbulence (2000 steps, kinetic energy and
                dissipation series), the closure network trained on its
                filtered snapshots, and 128^3 runs with and without it compared
                with the filtered 512^3 field
  reference_channel
                the reference's own problem (its 200x60x1 channel with a step,
                inflow/outflow) on the FP64 phase path, against the reference
                on one host core
  hbm_ceiling   what plain copy / read / write kernels reach on this GPU
  cpu_baseline  the unmodified reference (oracle/_ref, built from its own
                sources) under mpirun on this host's cores, bounded sample
  --calibrate-copy prints the HBM bytes a plain device copy of known size
  reports to rocprofv3's FETCH_SIZE/WRITE_SIZE counters (used to scale them).
"""
import argparse
import hashlib
import json
import os
import shutil
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ark-cpp-mpi_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell per step (each persistent array read once and
# written once): FP64 = (5 cell + 15 face fields) x 8 B x 2; mixed = 5 cell
# fields x 8 B x 2 + 15 face fields x 4 B x 2.
BYTES_PER_CELL = {"fp64": 320, "mixed": 200}
FUSED_KERNEL = {"fp64": "fused_step", "mixed": "fused_mixed"}
# Algorithmic HBM bytes per interior cell of the phase-by-phase FP64 kernels
# (every array element each kernel reads or writes, once): cfl reads u1..3;
# phase1 reads 5 level-n cell fields and 6 face fields per axis (u1,u2,u3,
# p, rho, t) and writes Con and nCon; stress_faces reads u1..3 and the
# condition and writes 9 stresses; forces reads condition, rho and 9 stresses
# and writes f1..3; use_forces reads condition, rho, u1..3, f1..3 and writes
# u1..3; phase2 (the three sweeps in one launch) reads 5 faces and 10 cell
# fields and writes 6 faces per axis; rho_faces reads and writes 3 face fields.
# A step's fused pairs: forces_use_forces reads condition, rho, rho_n, u1..3
# and 9 stresses and writes f1..3 and u1..3; phase1_use_forces is phase1 plus
# the reads of f1..3.
PHASE_BYTES_PER_CELL = {"cfl": 24, "phase1": 264, "stress_faces": 104, "forces": 112, "use_forces": 88,
                        "phase2": 3 * 168, "rho_faces": 48, "forces_use_forces": 168, "phase1_use_forces": 288}
# Ghost fills move planes: 3 cell fields x 3 axes x 2 sides before the stress;
# 7 cell fields and 5 face fields x 3 axes x 2 sides before Phase2 (read and
# write, 16 B per element).
GHOST_PLANES = {"ghost_stress": 3 * 3 * 2, "ghost_phase2": (7 + 5) * 3 * 2}


def time_steps(N, steps, warmup, precision, scheme, profile_steps):
    import arkamd
    prec = arkamd.PREC_MIXED if precision == "mixed" else arkamd.PREC_FP64
    s = arkamd.Solver(N, N, N, precision=prec, scheme=scheme)
    s.init_tgv(1.0)
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    diag = s.diagnostics(allow_nonfinite=True)
    prof = {}
    if profile_steps > 0:
        s.profile(True)
        s.profile_reset()
        s.advance(profile_steps)
        prof = s.profile_read()
        s.profile(False)
    dev_bytes = s.device_bytes()
    tile = ("30x14", "32x13")[s.tile_shape] if prec == arkamd.PREC_MIXED else "32x13"
    s.close()
    return elapsed, diag, prof, dev_bytes, tile


def roofline_of(prof, precision, cells):
    name = FUSED_KERNEL[precision]
    if name not in prof:
        return None
    launches, total_ms = prof[name]
    avg_s = total_ms / launches / 1e3
    algo = BYTES_PER_CELL[precision] * cells
    achieved = algo / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(name),
            "kernel": name, "algorithmic_bytes_per_launch": algo, "avg_launch_ms": round(avg_s * 1e3, 4),
            "launches_timed": launches}


def traffic_of(kernel):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
    (profiles/pmc_traffic.json: per-dispatch KiB scaled by what the counters
    read for a 1 GiB device copy), or None when no record exists or the
    kernel's source has changed since it was taken."""
    try:
        rec = json.load(open(os.path.join(REPO, "profiles", "pmc_traffic.json")))["kernels"][kernel]
        src = open(os.path.join(REPO, rec["source"]), "rb").read()
        if hashlib.sha256(src).hexdigest() != rec["source_sha256"]:
            return None
        return rec["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def phase_kernels(N, steps):
    """Achieved GB/s of every kernel of the phase-by-phase FP64 step."""
    import arkamd
    with arkamd.Solver(N, N, N) as s:
        s.init_tgv(1.0)
        s.step_phases(1)
        s.synchronize()
        s.profile(True)
        s.profile_reset()
        s.step_phases(steps)
        prof = s.profile_read()
    out = {}
    cells = N ** 3
    for name, (launches, total_ms) in sorted(prof.items()):
        avg_s = total_ms / launches / 1e3
        if name in PHASE_BYTES_PER_CELL:
            algo = PHASE_BYTES_PER_CELL[name] * cells
        elif name in GHOST_PLANES:
            algo = GHOST_PLANES[name] * N * N * 16
        else:
            algo = None
        rec = {"launches": launches, "avg_ms": round(avg_s * 1e3, 4)}
        if algo:
            rec.update({"bytes": algo, "GB/s": round(algo / avg_s / 1e9, 1),
                        "frac": round(algo / avg_s / 1e9 / HBM_PEAK_GBS, 4)})
        out[name] = rec
    return out


def batched(nboxes, steps, seed=4, width=4):
    """Many independent boxes in one launch per step: cubes with edges drawn
    from {32, 48, 64, 96, 128}, Re uniform in [400, 3200] (nu = U0 L / Re with
    the vortex scale L = 1), alternately a Taylor-Green vortex and a seeded
    solenoidal field (k0 = 4, u_rms 1) under random-phase forcing at eps = 0.1,
    mixed precision, corrected scheme, tile shape autotuned.  Aggregate
    Mcell-steps/s over all boxes; afterwards every box is box-filtered (width
    4 cells) into closure training pairs (filtered gradients and exact
    subgrid stresses)."""
    import numpy as np
    import arkamd
    rng = np.random.default_rng(seed)
    edges = rng.choice([32, 48, 64, 96, 128], size=nboxes)
    res = rng.uniform(400.0, 3200.0, size=nboxes)
    boxes = []
    for q, (e, r) in enumerate(zip(edges, res)):
        s = arkamd.Solver(int(e), int(e), int(e), VIS=1.0 / float(r), precision=arkamd.PREC_MIXED,
                          scheme=arkamd.SCHEME_CORRECTED)
        if q % 2:
            s.init_turbulence(1000 + q, 4.0, 1.0)
            s.set_forcing(2000 + q, 0.1, 0.01)
        else:
            s.init_tgv(1.0)
        boxes.append(s)
    cells = int(sum(int(e) ** 3 for e in edges))
    with arkamd.Batch(boxes) as b:
        tune = b.autotune(reps=3)
        cfg = b.config()
        b.step(2)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        el = time.perf_counter() - t0
    # the batch kernel alone (back-to-back launches of the kept configuration,
    # as the autotuner timed them) against the graph replay of whole steps
    # (the dt update, the two forcing launches and the batch kernel)
    kernel_ms = tune[(cfg["shape"], cfg["lz"])]
    nonfinite = sum(s.diagnostics(allow_nonfinite=True)["nonfinite"] for s in boxes)
    t0 = time.perf_counter()
    pairs = 0
    for s in boxes:
        ub, tau, grad = s.filter_box(width)
        pairs += ub[0].size
    filt = time.perf_counter() - t0
    for s in boxes:
        s.close()
    return {"boxes": nboxes, "cells": cells, "steps": steps, "edges_hist": {str(e): int((edges == e).sum())
                                                                             for e in (32, 48, 64, 96, 128)},
            "re_range": [round(float(res.min()), 1), round(float(res.max()), 1)],
            "kinds": {"taylor_green": (nboxes + 1) // 2, "forced_turbulence": nboxes // 2},
            "scheme": "corrected",
            "value": round(cells * steps / el / 1e6, 2), "unit": "Mcell-steps/s", "ms_per_step": round(el * 1e3 / steps, 4),
            "batch_kernel_ms": round(kernel_ms, 4),
            "outside_batch_kernel_frac": round((el * 1e3 / steps - kernel_ms) / (el * 1e3 / steps), 4),
            "config": cfg, "autotune_ms": {f"{k[0]}/lz{k[1]}": round(v, 4) for k, v in tune.items()},
            "nonfinite": nonfinite,
            "training_pairs": {"filter_width_cells": width, "pairs": pairs, "wall_s_incl_download": round(filt, 3)}}


def closure_experiment(N, steps, width, train_steps, batch, pairs=8, horizon=8):
    """Forced homogeneous turbulence at N^3 (the plain fine-grid CABARET step of
    the reference scheme, mixed precision): seeded solenoidal initial field
    E(k) ~ k^4 exp(-2 (k/4)^2), u_rms 1, random-phase forcing on 1 <= |k| <= 2
    at eps = 0.1 (redrawn every 0.01 in time), nu = 1/360 so that
    Re_lambda is about 100.  `pairs` times the filtered fine state (width w)
    starts a coarse run of `horizon` steps, whose lag behind the filtered fine
    run is the a-posteriori target of the closure network (the kept closure,
    its amplitude calibrated on the windows' start times: closure.calibrate);
    the exact subgrid stresses of the same states train a second network (the
    a-priori closure, for comparison).  (N/w)^3 runs from the filtered initial
    field, with and without each closure, are compared with the filtered fine
    run at the same time by the kinetic-energy spectrum and the dissipation."""
    from arkamd import closure as C
    vis, eps, seed = 1.0 / 360.0, 0.1, 31
    d = C.correction_pairs(N, steps, pairs, width, vis, eps, seed, horizon=horizon, series_every=100)
    tr = C.train(d["X"], d["T"], steps=train_steps, batch=batch, lr=1e-3, seed=1)
    ts = C.train(d["X"], d["T_sgs"], steps=train_steps, batch=batch, lr=1e-3, seed=1)
    Eref = C.spectrum_of(d["ub_end"], vis)
    nc = N // width
    delta = width * d["dx"]
    plain = C.les(d["ub0"], d["t_end"], vis, eps, seed)
    raw = C.les(d["ub0"], d["t_end"], vis, eps, seed, mlp=tr["mlp"], delta=delta)
    t0 = time.perf_counter()
    cal = C.calibrate(d, tr, vis, eps, seed, delta)
    cal["wall_s"] = round(time.perf_counter() - t0, 3)
    model = C.les(d["ub0"], d["t_end"], vis, eps, seed, mlp=tr["mlp"], delta=delta)
    prior = C.les(d["ub0"], d["t_end"], vis, eps, seed, mlp=ts["mlp"], delta=delta)
    tr["mlp"].close()
    ts["mlp"].close()
    last = d["series"][-1] if d["series"] else {}
    out = {"fine": {"grid": N, "steps": d["steps"], "time": d["t_end"], "wall_s": round(d["wall"], 3),
                    "Mcell_steps_per_s": round(N ** 3 * d["steps"] / d["wall"] / 1e6, 1), "nonfinite": d["nonfinite"],
                    "nu": vis, "dissipation_end": d["eps_end"], "Re_lambda_end": round(d["Re_lambda_end"], 1),
                    "eps_injected": eps,
                    "eps_balance": "not reached: dissipation %.3g against injection %.3g at t=%.3g (about half an "
                                   "eddy turnover; the longer runs of forced_turbulence_stationary reach "
                                   "it)" % (d["eps_end"], eps, d["t_end"])
                    if abs(d["eps_end"] - eps) > 0.1 * eps else "reached",
                    "series": [{k: (float("%.6g" % v) if isinstance(v, float) else v) for k, v in r.items()}
                               for r in d["series"]]},
           "training": {"pairs": int(d["X"].shape[0]), "snapshots": pairs, "horizon_coarse_steps": horizon,
                        "steps": train_steps, "batch": batch, "a_posteriori_loss": [tr["loss0"], tr["loss_end"]],
                        "a_priori_loss": [ts["loss0"], ts["loss_end"]], "wall_s": round(tr["wall"] + ts["wall"], 3)},
           "calibration": {"alpha": cal["alpha"], "wall_s": cal["wall_s"],
                           "candidates": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in t.items()}
                                          for t in cal["candidates"]],
                           "note": "amplitude of the a-posteriori closure, chosen by the mean spectrum error "
                                   "(shells 1..N/4) at the start times of the correction windows in the last "
                                   "quarter of the fine run; the final state takes no part"},
           "coarse_grid": nc, "filter_width_cells": width, "Re_lambda_last_series": last.get("Re_lambda")}
    runs = (("plain", plain), ("closure", model), ("closure_alpha_1", raw), ("closure_a_priori", prior))
    for name, r in runs:
        out[name] = {"steps": r["steps"], "wall_s": round(r["wall"], 4), "nonfinite": r["nonfinite"],
                     "spectrum_err_k<=%d" % (nc // 4): round(C.spectrum_error(r["E"], Eref, nc // 4), 4),
                     "spectrum_err_k<=%d" % (nc // 2): round(C.spectrum_error(r["E"], Eref, nc // 2), 4),
                     "dissipation_resolved": r["eps_resolved"],
                     "dissipation_err_vs_fine": round(abs(r["eps_resolved"] - d["eps_end"]) / d["eps_end"], 4),
                     "wall_ratio_fine_over_coarse": round(d["wall"] / r["wall"], 1)}
    out["spectrum_filtered_fine"] = [float("%.5g" % v) for v in Eref[: nc // 2 + 1]]
    for name, r in runs:
        out["spectrum_" + name] = [float("%.5g" % v) for v in r["E"][: nc // 2 + 1]]
    return out


def stationary_turbulence(t_end, grids=(64, 128)):
    """Forced homogeneous turbulence run to a statistically stationary state:
    the closure experiment's initial field, forcing and viscosity (eps = 0.1
    injected, random phases redrawn every 0.01, nu = 1/360), mixed precision,
    corrected scheme (the reference scheme's runs become non-finite near
    t = 14 at these grids).  Over the second half of the run: the kinetic
    energy per unit volume, its drift between the last two quarters, and the
    resolved viscous dissipation against the injected power.  With the energy
    stationary the total dissipation equals the injection; the resolved share
    rises with the grid, the rest being the scheme's own grid-scale
    dissipation."""
    import numpy as np
    import arkamd
    vis, eps, seed = 1.0 / 360.0, 0.1, 31
    out = {"eps_injected": eps, "nu": vis, "scheme": "corrected", "t_end": t_end}
    for N in grids:
        V = (2 * np.pi) ** 3
        with arkamd.Solver(N, N, N, VIS=vis, precision=arkamd.PREC_MIXED, scheme=arkamd.SCHEME_CORRECTED) as s:
            s.init_turbulence(seed, 4.0, 1.0)
            s.set_forcing(seed + 1, eps, 0.01)
            series, steps, nxt = [], 0, 0.0
            t0 = time.perf_counter()
            while True:
                d = s.diagnostics(allow_nonfinite=True)
                finite = bool(np.isfinite(d["ke_physical"]))
                if finite and d["time"] >= nxt:
                    series.append((d["time"], d["ke_physical"] / V, s.dissipation()))
                    nxt += 0.5
                if d["time"] >= t_end or not finite:
                    break
                s.step(100)
                steps += 100
            wall = time.perf_counter() - t0
        t = np.array([r[0] for r in series])
        ke = np.array([r[1] for r in series])
        er = np.array([r[2] for r in series])
        h, q3 = t >= t_end / 2, t >= 0.75 * t_end
        ke_mean, er_mean = float(ke[h].mean()), float(er[h].mean())
        up2 = 2.0 * ke_mean / 3.0
        out["%d^3" % N] = {
            "steps": steps, "wall_s": round(wall, 2), "finite": finite,
            "ke_per_volume_mean": round(ke_mean, 4), "ke_per_volume_range": [round(float(ke[h].min()), 4),
                                                                            round(float(ke[h].max()), 4)],
            "ke_drift_last_quarters": round(float((ke[q3].mean() - ke[h & ~q3].mean()) / ke_mean), 4),
            "eps_resolved_mean": round(er_mean, 5), "resolved_share_of_injection": round(er_mean / eps, 3),
            "Re_lambda_total_eps": round(float(up2 * np.sqrt(15.0 / (vis * eps))), 1),
            "Re_lambda_resolved_eps": round(float(up2 * np.sqrt(15.0 / (vis * er_mean))), 1),
            "series_every_1": [[round(a, 3), round(b, 5), round(c, 5)] for a, b, c in series[::2]]}
    return out


def les_timing(N, steps, width=4):
    """The learned-closure coarse step alone: an N^3 forced-turbulence run
    (nu = 1/360, mixed precision, reference scheme) with and without a closure
    network attached (weights from its seeded initialisation: the cost does
    not depend on them), ms per step of each, and the closure kernels' share
    (MLP inference, its inputs and the stress divergence) by HIP events."""
    import arkamd
    out = {}
    for name in ("plain", "closure"):
        s = arkamd.Solver(N, N, N, VIS=1.0 / 360.0, precision=arkamd.PREC_MIXED)
        s.init_turbulence(31, 4.0, 1.0)
        s.set_forcing(32, 0.1, 0.01)
        m = None
        if name == "closure":
            m = arkamd.MLP(seed=1)
            s.set_closure(m, width * s.dx[0])
        s.step(5)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        out[name] = {"ms_per_step": round((time.perf_counter() - t0) * 1e3 / steps, 4)}
        s.close()
        if m is not None:
            m.close()
    out["closure_overhead_ms_per_step"] = round(out["closure"]["ms_per_step"] - out["plain"]["ms_per_step"], 4)
    flop = 2 * (10 * 64 + 64 * 64 + 64 * 6) * N ** 3
    out["mlp_flop_per_step"] = flop
    out["grid"] = N
    out["steps"] = steps
    return out


STATE_FIELDS = ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
                "u11", "u21", "u31", "p1", "t1", "u12", "u22", "u32", "p2", "t2", "u13", "u23", "u33", "p3", "t3")


def pcie_inclusive(N, steps):
    """The mixed step with its state crossing the host boundary: the 20
    persistent arrays (5 cell, 15 face; FP64 host arrays with the reference's
    extents) downloaded and uploaded through ark_download / ark_upload from
    pageable host memory, timed apart from the steps.  Reported beside the
    headline, never as it: inputs resident in HBM are the headline's premise."""
    import time
    import arkamd
    with arkamd.Solver(N, N, N, precision=arkamd.PREC_MIXED) as s:
        s.init_tgv(1.0)
        s.step(2)
        s.synchronize()
        t0 = time.perf_counter()
        host = {n: s.download(n) for n in STATE_FIELDS}
        t_down = time.perf_counter() - t0
        t0 = time.perf_counter()
        for n in STATE_FIELDS:
            s.upload(n, host[n])
        s.synchronize()
        t_up = time.perf_counter() - t0
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        t_steps = time.perf_counter() - t0
    nbytes = sum(a.nbytes for a in host.values())
    cells = N ** 3
    return {"state_bytes": nbytes, "download_s": round(t_down, 4), "upload_s": round(t_up, 4),
            "download_GB_per_s": round(nbytes / t_down / 1e9, 2), "upload_GB_per_s": round(nbytes / t_up / 1e9, 2),
            "steps": steps, "steps_s": round(t_steps, 4),
            "Mcell_steps_per_s_state_in_and_out_once": round(cells * steps / (t_up + t_steps + t_down) / 1e6, 1),
            "Mcell_steps_per_s_state_in_and_out_every_step": round(cells / (t_up + t_down + t_steps / steps) / 1e6, 2)}


ACC_CELL = ["ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon"]
ACC_FACE = ["u11", "u21", "u31", "p1", "t1", "u12", "u22", "u32", "p2", "t2", "u13", "u23", "u33", "p3", "t3"]


def _interior(name, a, N):
    if name.endswith("Con"):
        return a[1:N + 1, 1:N + 1, 1:N + 1]
    d = int(name[-1]) - 1
    e = [N + (d == 0), N + (d == 1), N + (d == 2)]
    return a[1:e[2] + 1, 1:e[1] + 1, 1:e[0] + 1]


def _field_errors(mix, f64):
    """Relative L2 error of every persistent field, with a scale that stays
    meaningful where a field is near zero: a velocity component (cells or
    faces) against the norm of the whole FP64 velocity vector at the same
    positions (u3 of the vortex starts at zero and stays small); density,
    its perturbation rho - ro0 and the face pressures against themselves; t
    (constant on the vortex) as the largest absolute difference."""
    import numpy as np
    rel = lambda a, b: float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
    cell_u = np.sqrt(sum(np.sum(f64[n] ** 2) for n in ("u1nCon", "u2nCon", "u3nCon")))
    face_u = {d: np.sqrt(sum(np.sum(f64[f"u{m}{d}"] ** 2) for m in (1, 2, 3))) for d in (1, 2, 3)}
    err = {}
    for n in ACC_CELL + ACC_FACE:
        if n.startswith("u"):
            scale = cell_u if n.endswith("Con") else face_u[int(n[-1])]
            err[n] = float(np.sqrt(np.sum((mix[n] - f64[n]) ** 2)) / scale)
        elif n.startswith("t"):
            err[n] = float(np.abs(mix[n] - f64[n]).max())
        else:
            err[n] = rel(mix[n], f64[n])
    err["ronCon-ro0"] = rel(mix["ronCon"] - 1.0, f64["ronCon"] - 1.0)
    vel = ("u1nCon", "u2nCon", "u3nCon")
    err["velocity"] = float(np.sqrt(sum(np.sum((mix[n] - f64[n]) ** 2) for n in vel)) / cell_u)
    return err


def accuracy(N, steps):
    """Mixed against FP64 after `steps` steps, all 20 persistent fields, beside
    the FP32 envelope: the FP64 step started from the initial state rounded
    once to FP32, i.e. how far exact arithmetic itself moves each field when
    its input is perturbed by one FP32 rounding (what FP32 face storage does
    to every face every step)."""
    import numpy as np
    import arkamd
    runs = {}
    for key, prec, rounded in (("fp64", arkamd.PREC_FP64, False), ("mixed", arkamd.PREC_MIXED, False),
                               ("envelope", arkamd.PREC_FP64, True)):
        with arkamd.Solver(N, N, N, precision=prec, scheme=arkamd.SCHEME_CORRECTED) as s:
            s.init_tgv(1.0)
            if rounded:
                for n in ACC_CELL + ACC_FACE:
                    s.upload(n, s.download(n).astype(np.float32).astype(np.float64))
            s.step(steps)
            runs[key] = {n: _interior(n, s.download(n), N) for n in ACC_CELL + ACC_FACE}
    err = _field_errors(runs["mixed"], runs["fp64"])
    env = _field_errors(runs["envelope"], runs["fp64"])
    res = {"scheme": "corrected", "steps": steps,
           "bound": "rel L2 <= 1e-5 for the cell velocities and density",
           "velocity_rel_l2": err["velocity"], "ronCon_rel_l2": err["ronCon"],
           "fields": {n: {"err": err[n], "fp32_envelope": env[n],
                          "vs_envelope": round(err[n] / env[n], 3) if env[n] > 0 else None}
                      for n in err if n != "velocity"}}
    res["max_vs_envelope"] = max(v["vs_envelope"] for n, v in res["fields"].items()
                                 if v["vs_envelope"] is not None and not n.startswith("t"))
    return res


def cpu_baseline(N, steps, ranks):
    harness = os.path.join(REPO, "oracle", "_ref", "ark_ref_harness")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    if not (os.path.exists(harness) and os.path.exists(mpirun)):
        return None
    cmd = [mpirun, "-n", str(ranks), harness, "tgv", str(N), str(N), str(N), str(steps), "-"]
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=300, check=True).stdout
    except Exception as e:  # the baseline is reported, never required
        return {"error": str(e)[:200]}
    rec = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
    return {
        "value": round(rec["mcell_steps_per_s"], 3), "unit": "Mcell-steps/s", "cores": ranks, "kind": "reference",
        "sample": f"TGV {N}^3 x {steps} steps, FP64, mpirun -n {ranks} (dims {rec['dims']}), "
                  f"step loop only, max over ranks ({rec['loop_seconds']:.2f} s)",
    }


def reference_channel(steps, cpu_steps):
    """The reference's own problem (its Input()/InitializeData(): 200x60x1
    cells, backward-facing step, inflow/outflow, periods={0,0,0}) on the FP64
    phase path, bitwise equal to the reference, against the unmodified
    reference on one host core (the same harness, step loop only)."""
    import arkamd
    out = {"workload": "reference channel 200x60x1 (its Input/InitializeData), FP64, bitwise reference",
           "steps": steps}
    with arkamd.Solver.channel() as s:
        s.step(10)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        el = time.perf_counter() - t0
        d = s.diagnostics(allow_nonfinite=True)
    cells = 200 * 60 * 1
    out.update({"gpu_ms_per_step": round(el * 1e3 / steps, 4), "gpu_Mcell_steps_per_s": round(cells * steps / el / 1e6, 2),
                "nonfinite": d["nonfinite"], "time": d["time"]})
    harness = os.path.join(REPO, "oracle", "_ref", "ark_ref_harness")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    if cpu_steps > 0 and os.path.exists(harness) and os.path.exists(mpirun):
        tmp = os.path.join("/tmp", f"ark_channel_{os.getpid()}")
        try:
            res = subprocess.run([mpirun, "-n", "1", harness, "channel", str(cpu_steps), tmp], capture_output=True,
                                 text=True, timeout=300, check=True).stdout
            rec = json.loads([l for l in res.splitlines() if l.startswith("{")][-1])
            ms = rec["loop_seconds"] * 1e3 / cpu_steps
            out["cpu_reference"] = {"ms_per_step": round(ms, 3), "cores": 1, "steps": cpu_steps,
                                    "Mcell_steps_per_s": round(rec["mcell_steps_per_s"], 3)}
            out["speedup_vs_reference"] = round(ms / out["gpu_ms_per_step"], 1)
        except Exception as e:  # reported, never required
            out["cpu_reference"] = {"error": str(e)[:200]}
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    return out


def hbm_ceiling():
    """Best bandwidths of plain streaming kernels on this GPU (copy, read-only,
    write-only over 4 GiB; ark-cpp-mpi_amd/bin/ark_hbm_probe), the practical
    ceiling beside the published 8 TB/s; None if the probe is not built."""
    exe = os.path.join(REPO, "ark-cpp-mpi_amd", "bin", "ark_hbm_probe")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if out.returncode != 0:
        return None
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    rec["copy_frac_of_peak"] = round(max(rec["copy_GBs"], rec["copy_nt_GBs"]) / HBM_PEAK_GBS, 4)
    return rec


def calibrate_copy(nbytes):
    """Three plain device copies of nbytes (torch's elementwise copy kernel),
    whose FETCH_SIZE/WRITE_SIZE readings scale the counters to bytes."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=40,
                    help="fused steps timed per kernel with HIP events (32 or more: the FP64 record layout)")
    ap.add_argument("--accuracy-steps", type=int, default=500)
    ap.add_argument("--channel-steps", type=int, default=2000, help="steps of the reference's own channel (0: skip)")
    ap.add_argument("--closure-size", type=int, default=512, help="fine grid of the closure experiment (0: skip)")
    ap.add_argument("--closure-steps", type=int, default=2000)
    ap.add_argument("--closure-train-steps", type=int, default=20000)
    ap.add_argument("--batch-boxes", type=int, default=64, help="boxes of the batched run (0: skip)")
    ap.add_argument("--stationary-time", type=float, default=30.0,
                    help="time of the forced-turbulence runs to a stationary state, 64^3 and 128^3 (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--phase-steps", type=int, default=3,
                    help="steps of the phase-by-phase FP64 path timed per kernel (0: skip)")
    ap.add_argument("--pcie-steps", type=int, default=500,
                    help="steps of the run whose state crosses the host boundary (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=50,
                    help="steps of the CPU baseline's 256^3 sample (about 45 s on 16 host cores)")
    ap.add_argument("--cpu-ranks", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "mixed", "channel", "batched", "closure", "les", "stationary"], default=None,
                    help="time one precision only, or only the reference's channel (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one subdomain on one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import arkamd
    if args.only == "channel":
        print(json.dumps({"reference_channel": reference_channel(args.channel_steps, 0)}))
        return
    if args.only == "batched":
        print(json.dumps({"batched": batched(args.batch_boxes, args.batch_steps)}))
        return
    if args.only == "les":
        print(json.dumps({"les_timing": les_timing(args.closure_size // 4, args.steps)}))
        return
    if args.only == "stationary":
        print(json.dumps({"forced_turbulence_stationary": stationary_turbulence(args.stationary_time)}))
        return
    if args.only == "closure":
        print(json.dumps({"forced_turbulence_closure": closure_experiment(
            args.closure_size, args.closure_steps, 4, args.closure_train_steps, 65536)}))
        return
    N = args.size
    cells = N ** 3
    runs = {}
    for prec in ("mixed", "fp64"):
        if args.only and prec != args.only:
            continue
        elapsed, diag, prof, dev_bytes, tile = time_steps(N, args.steps, args.warmup, prec,
                                                          arkamd.SCHEME_REFERENCE, args.profile_steps)
        runs[prec] = {"elapsed": elapsed, "diag": diag, "prof": prof, "device_bytes": dev_bytes, "tile": tile,
                      "value": cells * args.steps / elapsed / 1e6, "ms_per_step": elapsed * 1e3 / args.steps,
                      "roofline": roofline_of(prof, prec, cells)}

    head = "mixed" if "mixed" in runs else "fp64"
    H = runs[head]
    line = {
        "metric": "Mcell-steps/s at 256^3 (FP64 vs mixed)",
        "value": round(H["value"], 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(H["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (mixed)" if head == "mixed" else "f64",
        "data": "synthetic (Taylor-Green vortex initial condition, generated on device)",
        "config": {"workload": f"Taylor-Green vortex {N}^3 periodic, Re=1600, c=10, CFL 0.3, "
                               f"{args.steps} steps, {head} precision, reference scheme",
                   "grid": [N, N, N], "precision": head, "tile_owned_columns": H["tile"]},
        "roofline": H["roofline"],
    }
    if "fp64" in runs and head != "fp64":
        F = runs["fp64"]
        line["fp64"] = {"value": round(F["value"], 2), "ms_per_step": round(F["ms_per_step"], 4),
                        "roofline": F["roofline"], "dtype": "f64",
                        "note": "bitwise equal to the reference step"}
        line["speedup_mixed_vs_fp64"] = round(H["value"] / F["value"], 3)
    line["kernels"] = {p: {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in r["prof"].items()}
                       for p, r in runs.items()}
    line["final"] = {p: {"dt": r["diag"]["dt"], "nstep": r["diag"]["nstep"], "nonfinite": r["diag"]["nonfinite"]}
                     for p, r in runs.items()}
    line["device_bytes"] = {p: r["device_bytes"] for p, r in runs.items()}
    if not args.only:
        line["hbm_ceiling"] = hbm_ceiling()
    if args.phase_steps > 0 and not args.only:
        line["phase_kernels_fp64"] = phase_kernels(N, args.phase_steps)
    if args.batch_boxes > 0 and not args.only:
        line["batched"] = batched(args.batch_boxes, args.batch_steps)
    if args.channel_steps > 0 and not args.only:
        line["reference_channel"] = reference_channel(args.channel_steps, 0 if args.no_cpu_baseline else 200)
    if args.closure_size > 0 and not args.only:
        line["forced_turbulence_closure"] = closure_experiment(args.closure_size, args.closure_steps, 4,
                                                               args.closure_train_steps, 65536)
    if args.stationary_time > 0 and not args.only:
        line["forced_turbulence_stationary"] = stationary_turbulence(args.stationary_time)
    if args.accuracy_steps > 0 and not args.only:
        line["accuracy"] = accuracy(N, args.accuracy_steps)
    if args.pcie_steps > 0 and not args.only:
        line["pcie_inclusive"] = pcie_inclusive(N, args.pcie_steps)
    base = None
    if not args.no_cpu_baseline:
        ranks = args.cpu_ranks or min(16, os.cpu_count() or 1)
        while ranks > 1 and (N % ranks):
            ranks -= 1
        base = cpu_baseline(N, args.cpu_steps, ranks)
    line["cpu_baseline"] = base
    print(json.dumps(line))


if __name__ == "__main__":
    main()
