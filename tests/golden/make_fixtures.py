"""Writes the golden vectors under tests/golden/ from the UNMODIFIED reference.

Runs only where the reference checkout and MPICH exist (the build container):
it drives ``oracle/_ref/ark_ref_harness`` (built by ``make -C oracle ref`` from
the reference's own sources) and stores inputs and outputs as small fixtures.
The reference itself never travels; only these data files do.

Fixtures written:
  phase_fluid.npz, phase_solid.npz
      a seeded random state on a non-cubic 6x5x4 box with distinct dx per
      axis (so a transposed index or axis mix-up fails), plus the reference's
      output of each phase function (TimeStepSize, Phase1, StressTensor,
      UseForces, Phase2) and of two full steps.  "solid" adds a random solid
      mask (condition=1 cells) and exercises the wall branches.
  tgv8.npz
      the full reference state (all 41 arrays with ghosts) of the 8^3
      Taylor-Green run after 1, 2 and 10 steps, with its dt/KE/mass series.
  bc.npz
      five seeded random cell arrays on a 5x4x3 box and the reference
      BoundaryConditions class's output (oracle/_ref/ark_ref_bc, built from
      source/BoundaryConditions.cpp and source/Arr3d.cpp) for alongX1/2/3,
      periodic and not.  ``python make_fixtures.py bc`` writes only this file.
  channel_digests.json
      the reference's own problem (its Input() and InitializeData(), the
      200x60x1 channel with a backward-facing step, periods={0,0,0}, one rank):
      per-array SHA-256 of all 41 arrays after 0, 1, 10, 100 and 200 steps and
      the dt/KE/mass series.  The run has its MPI receive buffers zero-filled
      (glibc MALLOC_PERTURB_=255, tcache off): with MPI_PROC_NULL neighbours
      the reference copies those uninitialised buffers into its ghost planes.
      Two more runs with other heap contents show which arrays depend on them
      (recorded as "buffer_dependence").  Also the SHA-256 of the reference's
      own output files for the channel (WriteDataParaView after 10 and 20
      steps, WriteAverageValues of the averages from step 5 to 20).
      ``python make_fixtures.py channel`` writes only this file;
      ``python make_fixtures.py channel_long`` adds the same run's digests
      after 2000 and 20000 steps ("long").
  tgv_digests.json
      per-array SHA-256 and norms of the 16^3 x 10-step and 64^3 x 200-step
      Taylor-Green runs (1 rank, full arrays), their per-step dt, KE and mass
      series, and the result of the decomposition-invariance check (the same
      32^3 run on 1, 2, 4 and 8 ranks must give bitwise-equal fields).
      ``python make_fixtures.py tgv256`` adds the 256^3 x 5 run (one rank, all
      arrays; eight ranks, interior cells); ``python make_fixtures.py
      tgv256_long`` the 256^3 x 50 run (eight ranks, interior cells).
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "oracle"))
from ark_oracle import FIELDS, OracleState, read_meta  # noqa: E402

HARNESS = os.path.join(REPO, "oracle", "_ref", "ark_ref_harness")
BC_HARNESS = os.path.join(REPO, "oracle", "_ref", "ark_ref_bc")
MPIRUN = "/opt/conda/bin/mpirun"
PHASES = ["TimeStepSize", "Phase1", "StressTensor", "UseForces", "Phase2", "Step"]


def run(args, ranks=1):
    cmd = [MPIRUN, "-n", str(ranks), HARNESS] + [str(a) for a in args]
    res = subprocess.run(cmd, check=True, capture_output=True, text=True)
    return res.stdout


def random_state(seed: int, solid: bool) -> OracleState:
    rng = np.random.default_rng(seed)
    N1, N2, N3 = 6, 5, 4
    st = OracleState(N1, N2, N3, dx=(0.3, 0.25, 0.2), dt=0.0125, CFL=0.3, VIS=0.05,
                     sound=10.0, ro0=1.0, t0=1.0)

    def vel(shape):
        v = rng.uniform(-1.0, 1.0, size=shape)
        v[rng.random(shape) < 0.12] = 0.0          # exact zeros: ucf/ucb == 0 branches
        v[rng.random(shape) < 0.05] *= -1e-3       # tiny values of both signs
        return v

    for name in FIELDS:
        shp = st.shape(name)
        if name.startswith("sigm") or name in ("f1", "f2", "f3"):
            st.arrays[name] = rng.uniform(-1, 1, size=shp)   # overwritten by the phases
        elif name == "condition":
            st.arrays[name] = (rng.random(shp) < 0.3).astype(np.float64) if solid else np.zeros(shp)
        elif name.startswith("ro"):
            st.arrays[name] = 1.0 + 0.02 * rng.uniform(-1, 1, size=shp)
        elif name.startswith("p"):
            st.arrays[name] = 2.0 * rng.uniform(-1, 1, size=shp)
        elif name.startswith("t"):
            st.arrays[name] = 1.0 + 0.1 * rng.uniform(-1, 1, size=shp)
        else:
            st.arrays[name] = vel(shp)
    # ties between a face value and its cell value hit the equal-operand paths of min3d/max3d
    st.arrays["u11"][2, 2, 3] = st.arrays["u1Con"][2, 2, 3]
    st.arrays["u22"][1, 3, 2] = st.arrays["u2Con"][1, 3, 2]
    return st


def reference_phase(state: OracleState, phase: str, work: str) -> OracleState:
    ind = os.path.join(work, "in")
    outd = os.path.join(work, "out_" + phase)
    shutil.rmtree(ind, ignore_errors=True)
    state.to_dir(ind)
    os.makedirs(outd, exist_ok=True)
    if phase == "Step":
        run(["run", ind, 2, outd])
    else:
        run(["phase", phase, ind, outd])
    return OracleState.from_dir(outd)


def pack_state(prefix: str, st: OracleState, names=None) -> dict:
    out = {f"{prefix}:{k}": np.asarray(v) for k, v in st.scalars().items()}
    for n in (names or FIELDS):
        out[f"{prefix}:{n}"] = st[n]
    return out


def phase_fixture(path: str, seed: int, solid: bool, work: str) -> None:
    st = random_state(seed, solid)
    data = pack_state("in", st)
    for ph in PHASES:
        res = reference_phase(st, ph, work)
        changed = [n for n in FIELDS if not np.array_equal(res[n].view(np.int64), st[n].view(np.int64))]
        data.update(pack_state(ph, res, changed))
        data[f"{ph}:changed"] = np.array(changed)
    np.savez_compressed(path, **data)


def digest(a: np.ndarray) -> dict:
    a = np.ascontiguousarray(a, dtype="<f8")
    flat = a.ravel()
    idx = np.linspace(0, flat.size - 1, 16).astype(np.int64)
    return {
        "sha256": hashlib.sha256(a.tobytes()).hexdigest(),
        "sum": float(flat.sum()).hex(), "l2": float(np.sqrt((flat * flat).sum())).hex(),
        "linf": float(np.abs(flat).max()).hex(), "min": float(flat.min()).hex(),
        "max": float(flat.max()).hex(),
        "sample_index": idx.tolist(), "sample": [float(flat[q]).hex() for q in idx],
    }


def read_series(path: str) -> dict:
    rows = [l.split() for l in open(os.path.join(path, "series.txt")) if not l.startswith("#")]
    return {"step": [int(r[0]) for r in rows], "dt": [r[1] for r in rows],
            "TIME": [r[2] for r in rows], "ke_reference": [r[3] for r in rows],
            "mass": [r[4] for r in rows]}


def tgv_run(N: int, steps: int, work: str, dumps, ranks=1) -> str:
    d = os.path.join(work, f"tgv{N}_{steps}_{ranks}")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    run(["-quiet", "tgv", N, N, N, steps, d] + list(dumps), ranks=ranks)
    return d


def bc_fixture(path: str, work: str) -> None:
    """Reference BoundaryConditions on seeded random arrays, every axis and flag."""
    N1, N2, N3 = 5, 4, 3
    t0 = 1.25
    rng = np.random.default_rng(777)
    inp = rng.uniform(-1.0, 2.0, size=(5, N3 + 2, N2 + 2, N1 + 2))
    fin = os.path.join(work, "bc_in.f64")
    inp.astype("<f8").tofile(fin)
    data = {"input": inp, "N": np.array([N1, N2, N3]), "t0": np.array(t0)}
    for axis in range(3):
        for periodic in (0, 1):
            fout = os.path.join(work, f"bc_{axis}_{periodic}.f64")
            subprocess.run([BC_HARNESS, str(N1), str(N2), str(N3), str(axis), str(periodic), repr(t0), fin, fout],
                           check=True)
            data[f"out_{axis}_{periodic}"] = np.fromfile(fout, dtype="<f8").reshape(inp.shape)
    np.savez_compressed(path, **data)


CHANNEL_ZERO_ENV = {"MALLOC_PERTURB_": "255", "GLIBC_TUNABLES": "glibc.malloc.tcache_count=0"}


def channel_run(work: str, steps: int, dumps, env_extra) -> str:
    d = os.path.join(work, "channel_" + "_".join(f"{k}{v}" for k, v in sorted(env_extra.items())).replace("=", "")
                     .replace(".", ""))
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    env = dict(os.environ)
    for k in ("MALLOC_PERTURB_", "GLIBC_TUNABLES"):
        env.pop(k, None)
    env.update(env_extra)
    subprocess.run([MPIRUN, "-n", "1", HARNESS, "channel", str(steps), d] + [str(x) for x in dumps], check=True,
                   capture_output=True, text=True, env=env)
    return d


def channel_fixture(path: str, work: str) -> None:
    steps, dumps = 200, [0, 1, 10, 100, 200]
    ref = channel_run(work, steps, dumps, CHANNEL_ZERO_ENV)
    out = {"problem": "reference Input()/InitializeData(): 201x61x2 nodes, 20 x 6 x 0.1, step x<=5 y<=1, "
                      "u10=1, c=10, VIS=1e-3, CFL=0.2, periods={0,0,0}, 1 rank",
           "N": [200, 60, 1], "env": CHANNEL_ZERO_ENV, "series": read_series(ref), "states": {}}
    for s in dumps:
        st = OracleState.from_dir(os.path.join(ref, f"s{s}"))
        out["states"][str(s)] = {n: digest(st[n]) for n in FIELDS}
    # which arrays depend on the uninitialised receive buffers' contents
    dep = {}
    for env_extra in ({}, {"MALLOC_PERTURB_": "165"}):
        other = channel_run(work, steps, [steps], env_extra)
        a = OracleState.from_dir(os.path.join(ref, f"s{steps}"))
        b = OracleState.from_dir(os.path.join(other, f"s{steps}"))
        N1, N2, N3 = a.N
        for n in FIELDS:
            x, y = a[n], b[n]
            e = 1 if n in FIELDS[FIELDS.index("ro1"):] else 0
            interior = (slice(1, N3 + 1 + e), slice(1, N2 + 1 + e), slice(1, N1 + 1 + e))
            dep.setdefault(n, {"whole_array": True, "interior": True})
            if not np.array_equal(x.view(np.uint64), y.view(np.uint64)):
                dep[n]["whole_array"] = False
            if not np.array_equal(x[interior].view(np.uint64), y[interior].view(np.uint64)):
                dep[n]["interior"] = False
    # the reference's own output files: WriteDataParaView at steps 10 and 20,
    # AverageValues from step 5 on and WriteAverageValues after step 20
    vd = os.path.join(work, "channel_vtk")
    shutil.rmtree(vd, ignore_errors=True)
    os.makedirs(vd)
    env = dict(os.environ)
    env.update(CHANNEL_ZERO_ENV)
    subprocess.run([MPIRUN, "-n", "1", HARNESS, "-vtk", "-avg", "5", "channel", "20", vd, "10", "20"], check=True,
                   capture_output=True, text=True, env=env)
    out["vtk"] = {"command": "ark_ref_harness -vtk -avg 5 channel 20 OUT 10 20", "files": {}}
    for fn in ("out_10.vtk", "out_20.vtk", "average.vtk"):
        data = open(os.path.join(vd, fn), "rb").read()
        out["vtk"]["files"][fn] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    out["buffer_dependence"] = {"independent_whole_array": sorted(n for n, v in dep.items() if v["whole_array"]),
                                "independent_interior": sorted(n for n, v in dep.items() if v["interior"]),
                                "dependent_interior": sorted(n for n, v in dep.items() if not v["interior"])}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


def channel_long_fixture(path: str, work: str, dumps=(2000, 20000)) -> None:
    """The channel run of channel_fixture continued to `dumps` steps: per-array
    SHA-256 of all 41 arrays and dt, TIME, KE and mass there, added to
    channel_digests.json as "long" (the full per-step series would be
    megabytes)."""
    dig = json.load(open(path))
    steps = max(dumps)
    ref = channel_run(work, steps, list(dumps), CHANNEL_ZERO_ENV)
    ser = read_series(ref)
    long = {"steps": list(dumps), "states": {}, "series_at": {}}
    for s in dumps:
        st = OracleState.from_dir(os.path.join(ref, f"s{s}"))
        long["states"][str(s)] = {n: digest(st[n]) for n in FIELDS}
        long["series_at"][str(s)] = {k: ser[k][s] for k in ser if k != "step"}
    dig["long"] = long
    with open(path, "w") as fh:
        json.dump(dig, fh, indent=1)


def tgv256_fixture(path: str, work: str, steps: int = 5) -> None:
    """The 256^3 Taylor-Green case of the GPU workload, `steps` steps of the
    unmodified reference: one rank (all 41 arrays with their ghosts, so the
    GPU's full state can be compared) and eight ranks (dims 4x2x1, z unsplit;
    the interior cell fields must be the same bits).  Adds the run to
    tgv_digests.json as "256^3x<steps>".  The eight-rank gather sums each
    cell with zeros from the other ranks, which turns -0.0 into +0.0, so its
    interior digests are taken after adding +0.0 on both sides."""
    dig = json.load(open(path))
    N = 256
    d = tgv_run(N, steps, work, [steps])
    sd = os.path.join(d, f"s{steps}")
    st = OracleState.from_dir(sd)
    meta = read_meta(sd)
    run1 = {
        "N": N, "steps": steps, "VIS": float(st.VIS).hex(), "CFL": float(st.CFL).hex(),
        "sound": float(st.sound).hex(), "final_dt": float(meta["dt"]).hex(),
        "series": read_series(d),
        "arrays": {n: digest(st[n]) for n in FIELDS},
    }
    interior = {n: np.ascontiguousarray(st[n][1:-1, 1:-1, 1:-1]) + 0.0
                for n in ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon")}
    del st
    shutil.rmtree(d, ignore_errors=True)
    h1 = {n: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
          for n, a in interior.items()}
    d8 = tgv_run(N, steps, work, [steps], ranks=8)
    h8 = {}
    for n in interior:
        a = np.fromfile(os.path.join(d8, f"s{steps}", n + ".interior.f64"), dtype="<f8").reshape(N, N, N) + 0.0
        h8[n] = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    ser8 = read_series(d8)
    shutil.rmtree(d8, ignore_errors=True)
    phys = [n for n in h1 if n != "tnCon"]
    if any(h1[n] != h8[n] for n in phys) or ser8["dt"] != run1["series"]["dt"]:
        raise SystemExit("256^3: eight ranks differ from one rank")
    run1["interior_sha256_plus_zero"] = h1
    run1["interior_sha256_plus_zero_8ranks"] = h8
    dig["runs"][f"{N}^3x{steps}"] = run1
    with open(path, "w") as fh:
        json.dump(dig, fh, indent=1)


def tgv256_long_fixture(path: str, work: str, steps: int = 50) -> None:
    """The 256^3 Taylor-Green case over `steps` steps of the unmodified
    reference on eight ranks (dims 4x2x1, z unsplit): SHA-256 of the gathered
    interior of the level-n cell fields (after adding +0.0, as in
    tgv256_fixture) and the dt series, added to tgv_digests.json as
    "256^3x<steps>_interior"."""
    dig = json.load(open(path))
    N = 256
    d8 = tgv_run(N, steps, work, [steps], ranks=8)
    h8 = {}
    for n in ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon"):
        a = np.fromfile(os.path.join(d8, f"s{steps}", n + ".interior.f64"), dtype="<f8").reshape(N, N, N) + 0.0
        h8[n] = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    dig["runs"][f"{N}^3x{steps}_interior"] = {"N": N, "steps": steps, "ranks": 8, "series": read_series(d8),
                                              "interior_sha256_plus_zero": h8}
    shutil.rmtree(d8, ignore_errors=True)
    with open(path, "w") as fh:
        json.dump(dig, fh, indent=1)


def main() -> None:
    if len(sys.argv) > 1 and sys.argv[1] == "tgv256_long":
        if not os.path.exists(HARNESS):
            sys.exit("reference harness not built: make -C oracle ref")
        work = tempfile.mkdtemp(prefix="ark_golden_")
        try:
            tgv256_long_fixture(os.path.join(HERE, "tgv_digests.json"), work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        print("256^3 x 50 interior digests added to", os.path.join(HERE, "tgv_digests.json"))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "tgv256":
        if not os.path.exists(HARNESS):
            sys.exit("reference harness not built: make -C oracle ref")
        work = tempfile.mkdtemp(prefix="ark_golden_")
        try:
            tgv256_fixture(os.path.join(HERE, "tgv_digests.json"), work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        print("256^3 digests added to", os.path.join(HERE, "tgv_digests.json"))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "channel_long":
        if not os.path.exists(HARNESS):
            sys.exit("reference harness not built: make -C oracle ref")
        work = tempfile.mkdtemp(prefix="ark_golden_")
        try:
            channel_long_fixture(os.path.join(HERE, "channel_digests.json"), work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        print("long channel digests added to", os.path.join(HERE, "channel_digests.json"))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "channel":
        if not os.path.exists(HARNESS):
            sys.exit("reference harness not built: make -C oracle ref")
        work = tempfile.mkdtemp(prefix="ark_golden_")
        try:
            channel_fixture(os.path.join(HERE, "channel_digests.json"), work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "bc":
        if not os.path.exists(BC_HARNESS):
            sys.exit("reference BoundaryConditions harness not built: make -C oracle ref")
        work = tempfile.mkdtemp(prefix="ark_golden_")
        try:
            bc_fixture(os.path.join(HERE, "bc.npz"), work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        print("bc fixture written to", HERE)
        return
    if not os.path.exists(HARNESS):
        sys.exit("reference harness not built: make -C oracle ref")
    work = tempfile.mkdtemp(prefix="ark_golden_")
    try:
        bc_fixture(os.path.join(HERE, "bc.npz"), work)
        phase_fixture(os.path.join(HERE, "phase_fluid.npz"), 1234, False, work)
        phase_fixture(os.path.join(HERE, "phase_solid.npz"), 4321, True, work)

        d8 = tgv_run(8, 10, work, [1, 2, 10])
        data = {}
        for s in (1, 2, 10):
            st = OracleState.from_dir(os.path.join(d8, f"s{s}"))
            data.update(pack_state(f"s{s}", st))
        ser = read_series(d8)
        data["series_dt"] = np.array([float.fromhex(v) for v in ser["dt"]])
        data["series_ke"] = np.array([float.fromhex(v) for v in ser["ke_reference"]])
        data["series_mass"] = np.array([float.fromhex(v) for v in ser["mass"]])
        np.savez_compressed(os.path.join(HERE, "tgv8.npz"), **data)

        dig = {"runs": {}}
        for N, steps in ((16, 10), (64, 200)):
            d = tgv_run(N, steps, work, [steps])
            st = OracleState.from_dir(os.path.join(d, f"s{steps}"))
            meta = read_meta(os.path.join(d, f"s{steps}"))
            dig["runs"][f"{N}^3x{steps}"] = {
                "N": N, "steps": steps, "VIS": float(st.VIS).hex(), "CFL": float(st.CFL).hex(),
                "sound": float(st.sound).hex(), "final_dt": float(meta["dt"]).hex(),
                "series": read_series(d),
                "arrays": {n: digest(st[n]) for n in FIELDS},
            }
        # decomposition invariance: 32^3 x 20 steps on 1, 2, 4, 8 ranks
        inv = {}
        base = None
        for ranks in (1, 2, 4, 8):
            d = tgv_run(32, 20, work, [20], ranks=ranks)
            sd = os.path.join(d, "s20")
            if ranks == 1:
                st = OracleState.from_dir(sd)
                fields = {n: np.ascontiguousarray(st[n][1:-1, 1:-1, 1:-1]) for n in
                          ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon")}
            else:
                fields = {n: np.fromfile(os.path.join(sd, n + ".interior.f64"), dtype="<f8").reshape(32, 32, 32)
                          for n in ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon")}
            h = {n: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
                 for n, a in fields.items()}
            inv[str(ranks)] = h
            # rho and u must agree bit for bit.  t may not: the reference never
            # exchanges t ghost faces (source/main.cpp:1101-1127), so at an
            # internal rank boundary its ghost t face keeps its initial value.
            phys = {n: v for n, v in h.items() if n != "tnCon"}
            if base is None:
                base = phys
            elif phys != base:
                raise SystemExit(f"decomposition invariance FAILED at {ranks} ranks")
        dig["decomposition_invariance_32^3x20"] = inv
        dig["decomposition_invariance_note"] = (
            "ronCon/u1nCon/u2nCon/u3nCon are bitwise equal on 1, 2, 4 and 8 ranks; tnCon may "
            "differ at round-off because t ghost faces are never exchanged by the reference")
        dig["generator"] = "tests/golden/make_fixtures.py via oracle/_ref/ark_ref_harness"
        with open(os.path.join(HERE, "tgv_digests.json"), "w") as fh:
            json.dump(dig, fh, indent=1)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
