#!/usr/bin/env python3
"""Benchmark of the MI355X CABARET/ARK step: Mcell-steps/s on the 256^3
Taylor-Green vortex (periodic box of side 2 pi, c=10, Re=1600, CFL 0.3).

One "step" is one iteration of the reference's driver loop
(source/main.cpp:82-91) over the whole box, inputs resident in HBM.
Prints ONE JSON line.  Extra objects:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over its
                average launch time (HIP events on the library's own stream)
  cpu_baseline  the unmodified reference (oracle/_ref, built from the
                reference's sources) under mpirun on this host's cores, on a
                bounded sample of the same workload
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ark-cpp-mpi_amd"))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of work for each kernel: every array element
# the kernel must read or write, counted once.  Units are cells, or faces of
# one orientation for the Phase2 sweeps.
def kernel_bytes(prec):
    d = 8  # FP64 storage
    return {
        # fused predictor: reads 5 level-n cell fields + 15 face fields (u1,u2,u3,p,t
        # per axis), writes 5 half-level fields and 3 forces
        "fused_predictor": (5 + 15 + 5 + 3) * d,
        # fused corrector: reads 5 level-n + 5 half-level cell fields, 15 old faces,
        # 3 forces; writes 15 new faces and 5 level-(n+1) cell fields
        "fused_corrector": (5 + 5 + 15 + 3 + 15 + 5) * d,
        "phase1": (5 + 18 + 5 + 5) * d,
        "stress_faces": (3 + 9) * d,
        "forces": (9 + 1 + 3) * d,
        "use_forces": (4 + 3 + 3) * d,
        "phase2_x": (10 + 5 + 6) * d,
        "phase2_y": (10 + 5 + 6) * d,
        "phase2_z": (10 + 5 + 6) * d,
        "cfl": 3 * d,
    }


def run_gpu(N, steps, warmup, precision, profile_steps):
    import arkamd
    s = arkamd.Solver(N, N, N, precision=precision)
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
        s.step(profile_steps)
        prof = s.profile_read()
        s.profile(False)
    dev_bytes = s.device_bytes()
    s.close()
    return elapsed, diag, prof, dev_bytes


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
        "value": rec["mcell_steps_per_s"], "unit": "Mcell-steps/s", "cores": ranks, "kind": "reference",
        "sample": f"TGV {N}^3 x {steps} steps, FP64, mpirun -n {ranks} (dims {rec['dims']}), "
                  f"step loop only, max over ranks ({rec['loop_seconds']:.2f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-ranks", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one subdomain on one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import arkamd
    prec = arkamd.PREC_MIXED if args.precision == "mixed" else arkamd.PREC_FP64
    N = args.size
    cells = N ** 3
    elapsed, diag, prof, dev_bytes = run_gpu(N, args.steps, args.warmup, prec, args.profile_steps)
    ms_per_step = elapsed * 1e3 / args.steps
    value = cells * args.steps / elapsed / 1e6

    roofline = None
    kernels = {}
    if prof:
        kb = kernel_bytes(prec)
        for name, (launches, total_ms) in prof.items():
            kernels[name] = {"launches": launches, "avg_ms": total_ms / launches}
        dom = max(prof.items(), key=lambda kv: kv[1][1])
        name, (launches, total_ms) = dom
        avg_s = total_ms / launches / 1e3
        if name in kb:
            unit_count = cells
            algo = kb[name] * unit_count
            achieved = algo / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
                        "algorithmic_bytes_per_launch": algo, "avg_launch_ms": round(avg_s * 1e3, 4)}

    base = None
    if not args.no_cpu_baseline:
        ranks = args.cpu_ranks or min(16, os.cpu_count() or 1)
        while ranks > 1 and (N % ranks):
            ranks -= 1
        base = cpu_baseline(N, args.cpu_steps, ranks)

    line = {
        "metric": "Mcell-steps/s at 256^3 (FP64 vs mixed)",
        "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == arkamd.PREC_FP64 else "f32+f64",
        "data": "synthetic (Taylor-Green vortex initial condition, generated on device)",
        "config": {"workload": f"TGV {N}^3 periodic, Re=1600, c=10, CFL 0.3, {args.precision}",
                   "grid": [N, N, N], "precision": args.precision},
        "roofline": roofline,
        "cpu_baseline": base,
        "kernels": kernels,
        "final": {"dt": diag["dt"], "nstep": diag["nstep"], "ke_physical": diag["ke_physical"],
                  "nonfinite": diag["nonfinite"]},
        "device_bytes": dev_bytes,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
