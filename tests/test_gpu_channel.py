"""GPU: the reference's own problem, the channel with a backward-facing step
(its Input() and InitializeData(), source/main.cpp:118-372), on one rank with
periods={0,0,0}: inflow and outflow on X1, zero-gradient ends on X1 and X3,
walls and the step as solid cells.  The FP64 phase path must reproduce the
reference's run bit for bit on all 41 arrays, ghosts included, against the
digests the unmodified reference wrote (tests/golden/channel_digests.json)
and against the oracle step by step."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits_equal, describe_mismatch
from ark_oracle import FIELDS, OracleState
import arkamd

pytestmark = pytest.mark.gpu


def channel_solver(st: OracleState) -> arkamd.Solver:
    s = arkamd.Solver(*st.N, dx=st.dx, CFL=st.CFL, VIS=st.VIS, sound=st.sound, ro0=st.ro0, t0=st.t0,
                      periodic=(0, 0, 0))
    s.upload_all(st.arrays)
    return s


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_channel_matches_reference_digests():
    dig = json.load(open(os.path.join(GOLDEN, "channel_digests.json")))
    st = OracleState.channel()
    with channel_solver(st) as s:
        done = 0
        for k in sorted(int(x) for x in dig["states"]):
            if k > done:
                s.step_phases(k - done)
                done = k
            got = s.download_all()
            bad = [n for n in FIELDS if sha(got[n]) != dig["states"][str(k)][n]["sha256"]]
            assert not bad, (k, bad)
        assert s.diagnostics()["nstep"] == done
        assert s.dt == float.fromhex(dig["series"]["dt"][done])


def test_channel_long_run_matches_reference_digests():
    """The same channel run continued to 2000 and 20000 steps (TIME 3.6 and
    35.3; the reference harness took 13 minutes for it on one host core, the
    GPU takes about a second): all 41 arrays bitwise against the unmodified
    reference's digests there, dt and TIME bitwise, and its kinetic-energy
    and mass sums to 1e-13 (the library sums in a fixed tree order, the
    reference in loop order).  The reference's kinetic energy (WriteEnergy,
    source/main.cpp:2851-2876) is the sum of |u|^2/2 times the domain volume,
    20 x 6 x 0.1 here; the library reports the sum alone."""
    dig = json.load(open(os.path.join(GOLDEN, "channel_digests.json")))["long"]
    st = OracleState.channel()
    volume = float(np.prod([n * h for n, h in zip(st.N, st.dx)]))
    with channel_solver(st) as s:
        done = 0
        for k in dig["steps"]:
            s.step(k - done)
            done = k
            got = s.download_all()
            bad = [n for n in FIELDS if sha(got[n]) != dig["states"][str(k)][n]["sha256"]]
            assert not bad, (k, bad)
            ref = dig["series_at"][str(k)]
            d = s.diagnostics()
            assert d["nstep"] == k and d["nonfinite"] == 0
            assert s.dt == float.fromhex(ref["dt"]) and d["time"] == float.fromhex(ref["TIME"]), k
            assert d["ke_reference"] * volume == pytest.approx(float.fromhex(ref["ke_reference"]), rel=1e-13)
            assert d["mass"] == pytest.approx(float.fromhex(ref["mass"]), rel=1e-13)


def test_channel_phase_by_phase_vs_oracle():
    """Each reference phase on its own, first step and step 20, against the
    oracle, every array bitwise (localises a mismatch to one phase)."""
    ref = OracleState.channel()
    with channel_solver(ref) as s:
        for step in range(1, 21):
            for ph in ("TimeStepSize", "Phase1", "StressTensor", "UseForces", "Phase2", "Phase1", "UseForces"):
                if ph == "TimeStepSize":
                    ref.dt = ref.time_step_size()
                    assert s.time_step_size() == ref.dt
                else:
                    ref.phase(ph)
                    s.phase(ph)
                if step in (1, 20):
                    got = s.download_all()
                    bad = [describe_mismatch(n, got[n], ref[n]) for n in FIELDS if not bits_equal(got[n], ref[n])]
                    assert not bad, (step, ph, bad)


def test_channel_step_routes_to_the_phase_path():
    """ark_step on a non-periodic box runs the phase path (the fused kernel
    needs a periodic fluid box), with the same result; ark_advance and mixed
    precision refuse it."""
    st = OracleState.channel()
    with channel_solver(st) as a, channel_solver(st) as b:
        a.step(7)
        b.step_phases(7)
        ga, gb = a.download_all(), b.download_all()
        assert all(bits_equal(ga[n], gb[n]) for n in FIELDS)
        with pytest.raises(arkamd.ArkError):
            a.advance(1)
    with pytest.raises(arkamd.ArkError):
        arkamd.Solver(*st.N, dx=st.dx, precision=arkamd.PREC_MIXED, periodic=(0, 0, 0))


def test_solid_cells_on_a_periodic_box_vs_oracle():
    """The phase path with solid cells on a periodic box (the wall branches of
    every sweep, including X2's), from the seeded solid fixture state."""
    from ark_oracle import state_from_npz
    z = np.load(os.path.join(GOLDEN, "phase_solid.npz"))
    st = state_from_npz(z, "in")
    ref = st.copy()
    with arkamd.Solver(*st.N, dx=st.dx, CFL=st.CFL, VIS=st.VIS, sound=st.sound, ro0=st.ro0, t0=st.t0) as s:
        s.upload_all(st.arrays)
        s.dt = st.dt
        ref.step(3)
        s.step_phases(3)
        got = s.download_all()
        bad = [describe_mismatch(n, got[n], ref[n]) for n in FIELDS if not bits_equal(got[n], ref[n])]
        assert not bad, bad


def test_channel_output_files_match_reference(tmp_path):
    """The reference's output path, byte for byte: WriteDataParaView after 10
    and 20 steps (velocity, density, vorticity from the faces, big-endian)
    and WriteAverageValues of the running averages from step 5 to step 20."""
    files = json.load(open(os.path.join(GOLDEN, "channel_digests.json")))["vtk"]["files"]
    st = OracleState.channel()
    with channel_solver(st) as s:
        start = None
        for n in range(1, 21):
            s.step_phases(1)
            if n >= 5:
                if n == 5:
                    start = s.diagnostics()["time"]
                s.average_values()
            if n in (10, 20):
                path = s.write_paraview(str(tmp_path))
                assert os.path.basename(path) == f"out_{n}.vtk"
                data = open(path, "rb").read()
                assert len(data) == files[f"out_{n}.vtk"]["bytes"]
                assert hashlib.sha256(data).hexdigest() == files[f"out_{n}.vtk"]["sha256"], n
        data = open(s.write_average_values(str(tmp_path), start), "rb").read()
        assert len(data) == files["average.vtk"]["bytes"]
        assert hashlib.sha256(data).hexdigest() == files["average.vtk"]["sha256"]


def test_library_channel_init_matches_reference():
    """ark_init_channel (the product's restatement of InitializeData) gives the
    reference's initial arrays bit for bit, and runs to its step-10 state."""
    dig = json.load(open(os.path.join(GOLDEN, "channel_digests.json")))
    with arkamd.Solver.channel() as s:
        for k in (0, 10):
            if k:
                s.step(k)
            got = s.download_all()
            bad = [n for n in FIELDS if sha(got[n]) != dig["states"][str(k)][n]["sha256"]]
            assert not bad, (k, bad)


def test_channel_graph_steps_equal_single_steps():
    """A long ark_step on the phase path replays pairs of steps from one
    captured graph; its results, read between calls of several lengths (odd
    ones included), equal single phase steps bit for bit."""
    with arkamd.Solver.channel() as a, arkamd.Solver.channel() as b:
        done = 0
        for n in (17, 9, 1, 30, 8):
            a.step(n)
            b.step_phases(n)
            done += n
            ga, gb = a.download_all(), b.download_all()
            bad = [describe_mismatch(f, ga[f], gb[f]) for f in FIELDS if not bits_equal(ga[f], gb[f])]
            assert not bad, (done, bad)
            da, db = a.diagnostics(allow_nonfinite=True), b.diagnostics(allow_nonfinite=True)
            assert da["dt"] == db["dt"] and da["time"] == db["time"] and da["nstep"] == db["nstep"] == done
