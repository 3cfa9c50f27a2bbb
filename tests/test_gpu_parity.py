"""HIP step against the oracle and the reference's golden vectors.

FP64 mode is bit-for-bit: every comparison below is on the raw float64 bits of
every one of the reference's 41 arrays, ghosts included, unless it says
otherwise.  All calls go through the C-ABI (lib/libark_amd.so) via ctypes.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits_equal, describe_mismatch
from ark_oracle import FIELDS, OracleState, state_from_npz
import arkamd

pytestmark = pytest.mark.gpu

PHASES = ["TimeStepSize", "Phase1", "StressTensor", "UseForces", "Phase2"]


def solver_for(st: OracleState, precision=arkamd.PREC_FP64) -> arkamd.Solver:
    s = arkamd.Solver(*st.N, dx=st.dx, CFL=st.CFL, VIS=st.VIS, sound=st.sound, ro0=st.ro0, t0=st.t0,
                      precision=precision)
    s.upload_all(st.arrays)
    s.dt = st.dt
    return s


def assert_state_equal(solver: arkamd.Solver, want: OracleState, names=FIELDS, what=""):
    got = solver.download_all(names)
    bad = [describe_mismatch(n, got[n], want[n]) for n in names if not bits_equal(got[n], want[n])]
    assert not bad, what + "\n" + "\n".join(bad)


def random_fluid_state(N, seed, dt=0.01) -> OracleState:
    rng = np.random.default_rng(seed)
    st = OracleState(*N, dx=(0.31, 0.27, 0.22), dt=dt, CFL=0.3, VIS=0.02, sound=10.0, ro0=1.0, t0=1.0)
    for name in FIELDS:
        shp = st.shape(name)
        if name == "condition":
            continue
        if name.startswith("ro"):
            st.arrays[name] = 1.0 + 0.02 * rng.uniform(-1, 1, shp)
        elif name.startswith("p"):
            st.arrays[name] = 2.0 * rng.uniform(-1, 1, shp)
        elif name.startswith("t"):
            st.arrays[name] = 1.0 + 0.1 * rng.uniform(-1, 1, shp)
        else:
            v = rng.uniform(-1, 1, shp)
            v[rng.random(shp) < 0.1] = 0.0
            st.arrays[name] = v
    return st


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [(8, 8, 8), (6, 5, 4), (17, 9, 12), (12, 10, 8), (15, 20, 30)])
def test_init_tgv_bitwise(N):
    ref = OracleState.tgv(*N)
    with arkamd.Solver(*N) as s:
        s.init_tgv(1.0)
        assert_state_equal(s, ref, what=f"TGV init {N}")


@pytest.mark.parametrize("phase", PHASES)
def test_phase_vs_reference_fixture(phase):
    """Each phase on the reference's own random non-cubic input (tests/golden)."""
    z = np.load(os.path.join(GOLDEN, "phase_fluid.npz"))
    st = state_from_npz(z, "in")
    changed = set(str(n) for n in z[f"{phase}:changed"])
    want = st.copy()
    for n in changed:
        want.arrays[n] = z[f"{phase}:{n}"]
    with solver_for(st) as s:
        if phase == "TimeStepSize":
            dt = s.time_step_size()
            assert dt == float(z[f"{phase}:dt"])
        else:
            s.phase(phase)
        assert_state_equal(s, want, what=phase)


@pytest.mark.parametrize("stepper", ["step", "step_phases"])
def test_two_steps_vs_reference_fixture(stepper):
    z = np.load(os.path.join(GOLDEN, "phase_fluid.npz"))
    st = state_from_npz(z, "in")
    changed = set(str(n) for n in z["Step:changed"])
    want = st.copy()
    for n in changed:
        want.arrays[n] = z[f"Step:{n}"]
    with solver_for(st) as s:
        getattr(s, stepper)(2)
        assert s.dt == float(z["Step:dt"])
        assert_state_equal(s, want, what=stepper)


@pytest.mark.parametrize("N,seed", [((7, 3, 5), 11), ((16, 16, 16), 12), ((33, 17, 9), 13)])
@pytest.mark.parametrize("phase", PHASES)
def test_phase_vs_oracle_random(N, seed, phase):
    st = random_fluid_state(N, seed)
    want = st.copy()
    want.phase(phase)
    with solver_for(st) as s:
        if phase == "TimeStepSize":
            assert s.time_step_size() == want.dt
        else:
            s.phase(phase)
        assert_state_equal(s, want, what=f"{phase} {N}")


def test_tgv8_vs_reference_states():
    z = np.load(os.path.join(GOLDEN, "tgv8.npz"))
    with arkamd.Solver(8, 8, 8) as s:
        s.init_tgv()
        done = 0
        for target in (1, 2, 10):
            for _ in range(target - done):
                s.step(1)
            done = target
            assert_state_equal(s, state_from_npz(z, f"s{target}"), what=f"8^3 step {target}")


@pytest.mark.parametrize("N,steps", [(16, 10), (64, 200)])
def test_tgv_digest(N, steps):
    """The reference's 16^3 x 10 and 64^3 x 200 Taylor-Green runs: SHA-256 of
    all 41 final arrays and the whole dt series, bit for bit."""
    run = json.load(open(os.path.join(GOLDEN, "tgv_digests.json")))["runs"][f"{N}^3x{steps}"]
    want_dt = [float.fromhex(v) for v in run["series"]["dt"][1:]]
    with arkamd.Solver(N, N, N) as s:
        s.init_tgv()
        dts = []
        for _ in range(5):          # one call per step: the dt series, step by step
            s.step(1)
            dts.append(s.dt)
        assert dts == want_dt[:5]
        s.step(steps - 5)           # the rest in one call: fused kernels
        assert s.dt == want_dt[-1]
        got = s.download_all()
        bad = [n for n in FIELDS
               if hashlib.sha256(np.ascontiguousarray(got[n], dtype="<f8").tobytes()).hexdigest()
               != run["arrays"][n]["sha256"]]
        assert not bad, f"arrays differing from the reference digest: {bad}"
        d = s.diagnostics()
        assert d["nstep"] == steps


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("how", ["step", "advance"])
def test_tgv_digest_256(how):
    """The 256^3 workload against the unmodified reference (5 steps, one rank;
    its eight-rank run gives the same interior cells).  `step`: four steps
    through the production fused kernel (32x13 tiles, the chip-sized chunk
    table, 32-bit byte offsets up to the top of a field group) and the last
    through the phase kernels, then all 41 arrays, ghosts included, by
    SHA-256.  `advance`: five fused steps, then the interior of the level-n
    cell fields (after adding +0.0, as the reference's gathered digest) and
    the final dt."""
    run = json.load(open(os.path.join(GOLDEN, "tgv_digests.json")))["runs"]["256^3x5"]
    want_dt = [float.fromhex(v) for v in run["series"]["dt"][1:]]
    N = 256
    with arkamd.Solver(N, N, N) as s:
        s.init_tgv()
        if how == "step":
            s.step(5)
            assert s.dt == want_dt[-1]
            bad = [n for n in FIELDS if sha(s.download(n)) != run["arrays"][n]["sha256"]]
            assert not bad, f"arrays differing from the reference digest: {bad}"
        else:
            s.advance(4)
            assert s.time_step_size() == want_dt[4]
            s.advance(1)
            assert s.dt == want_dt[-1]
            for n, h in run["interior_sha256_plus_zero"].items():
                got = s.download(n)[1:-1, 1:-1, 1:-1] + 0.0
                assert sha(got) == h, n
        assert s.diagnostics()["nstep"] == 5


@pytest.mark.parametrize("how", ["step", "advance", "records"])
def test_tgv_digest_256_long(how):
    """Fifty steps of the 256^3 workload against the unmodified reference on
    eight ranks (dims 4x2x1): the dt of every tenth step and, at the end, the
    interior of the level-n density and velocities (after adding +0.0, as the
    reference's gathered digest).  The passive t is not compared: the
    reference never exchanges the t ghost faces (source/main.cpp:1047-1698),
    so on eight ranks its t differs from one rank's, which this GPU run is
    (the 5-step run above records both).  `advance`: fused steps only;
    `step`: each call of ten runs nine fused steps and one through the phase
    kernels; `records`: ark_step(50) in one call, 49 fused steps in the
    record layout, then one through the phase kernels (only its final dt is
    compared)."""
    run = json.load(open(os.path.join(GOLDEN, "tgv_digests.json")))["runs"]["256^3x50_interior"]
    want_dt = [float.fromhex(v) for v in run["series"]["dt"][1:]]
    assert len(want_dt) == 50
    N = 256
    with arkamd.Solver(N, N, N) as s:
        s.init_tgv()
        if how == "records":
            s.step(50)
            assert s.dt == want_dt[49]
        for done in range(10, 51, 10) if how != "records" else ():
            s.step(10) if how == "step" else s.advance(10)
            assert s.dt == want_dt[done - 1], done
        for n in ("ronCon", "u1nCon", "u2nCon", "u3nCon"):
            got = s.download(n)[1:-1, 1:-1, 1:-1] + 0.0
            assert sha(got) == run["interior_sha256_plus_zero"][n], n
        assert s.diagnostics()["nstep"] == 50


@pytest.mark.parametrize("N", [(32, 32, 32), (20, 13, 7), (70, 40, 90)])
def test_fused_equals_phases(N):
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        a.step(7)
        b.step_phases(7)
        ga, gb = a.download_all(), b.download_all()
        bad = [describe_mismatch(n, ga[n], gb[n]) for n in FIELDS if not bits_equal(ga[n], gb[n])]
        assert not bad, "\n".join(bad)
        da, db = a.diagnostics(), b.diagnostics()
        assert da["dt"] == db["dt"] and da["time"] == db["time"] and da["nstep"] == db["nstep"] == 7


@pytest.mark.parametrize("N", [(32, 32, 32), (20, 13, 7)])
def test_fused_equals_phases_record_layout(N):
    """A call with 32 or more fused steps first converts the four field groups
    to 32-byte records (fields 0-3 of an index; field 4 per field), both
    buffers index for index.  40 steps that way equal 40 phase-by-phase steps
    on every one of the 41 arrays."""
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        a.step(40)
        b.step_phases(40)
        ga, gb = a.download_all(), b.download_all()
        bad = [describe_mismatch(n, ga[n], gb[n]) for n in FIELDS if not bits_equal(ga[n], gb[n])]
        assert not bad, "\n".join(bad)
        da, db = a.diagnostics(), b.diagnostics()
        assert da["dt"] == db["dt"] and da["time"] == db["time"] and da["nstep"] == db["nstep"] == 40


def test_record_layout_is_exact_everywhere():
    """The record layout changes no bit anywhere: ark_advance in one run of 45
    (records) and in runs of 9 (per field) leave all 41 arrays equal, ghosts
    and the entries a fused step does not write included.  So do a following
    ark_step(36) (35 fused steps in records, after a download between the two
    record runs, then one step phase by phase) and the same 35 fused steps in
    runs below the threshold followed by ark_step(1)."""
    N = (20, 13, 7)
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        a.advance(45)
        for _ in range(5):
            b.advance(9)
        for what in ("advance", "step"):
            ga, gb = a.download_all(), b.download_all()
            bad = [describe_mismatch(n, ga[n], gb[n]) for n in FIELDS if not bits_equal(ga[n], gb[n])]
            assert not bad, what + "\n" + "\n".join(bad)
            if what == "advance":
                a.step(36)
                for k in (9, 9, 9, 8):
                    b.advance(k)
                b.step(1)
        da, db = a.diagnostics(), b.diagnostics()
        assert da["time"] == db["time"] and da["nstep"] == db["nstep"] == 81


def test_fused_equals_phases_random_state():
    st = random_fluid_state((19, 11, 6), 99, dt=0.004)
    with solver_for(st) as a, solver_for(st) as b:
        a.step(3)
        b.step_phases(3)
        ga, gb = a.download_all(), b.download_all()
        bad = [describe_mismatch(n, ga[n], gb[n]) for n in FIELDS if not bits_equal(ga[n], gb[n])]
        assert not bad, "\n".join(bad)


def test_determinism_two_runs():
    outs = []
    for _ in range(2):
        with arkamd.Solver(24, 24, 24) as s:
            s.init_tgv()
            s.step(5)
            d = s.diagnostics()
            outs.append((s.download_all(["u1nCon", "ronCon", "p2"]), d))
    for n in outs[0][0]:
        assert bits_equal(outs[0][0][n], outs[1][0][n])
    assert outs[0][1] == outs[1][1]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reinit_after_fused_steps_matches_fresh_context(prec):
    """Fused steps leave the next dt as a pending running MIN; a context
    initialised again afterwards must step exactly as a fresh one."""
    p = arkamd.PREC_FP64 if prec == "fp64" else arkamd.PREC_MIXED
    with arkamd.Solver(20, 16, 12, precision=p) as a, arkamd.Solver(20, 16, 12, precision=p) as b:
        a.init_tgv()
        a.advance(3)
        a.init_tgv()
        a.step(4)
        b.init_tgv()
        b.step(4)
        da, db = a.diagnostics(), b.diagnostics()
        assert da["dt"] == db["dt"] and da["nstep"] == db["nstep"]
        for n in ("u1nCon", "ronCon", "p1", "t2"):
            assert bits_equal(a.download(n), b.download(n)), n


def test_diagnostics_vs_oracle():
    st = OracleState.tgv(16)
    st.step(3)
    with arkamd.Solver(16, 16, 16) as s:
        s.init_tgv()
        s.step(3)
        d = s.diagnostics()
    assert d["ke_reference"] == pytest.approx(st.energy_sum(), rel=1e-13)
    assert d["mass"] == pytest.approx(st.mass(), rel=1e-13)
    u = [st[n][1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")]
    ro = st["ronCon"][1:-1, 1:-1, 1:-1]
    dvc = st.dx[0] * st.dx[1] * st.dx[2]
    kp = float((0.5 * ro * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2) * dvc).sum())
    assert d["ke_physical"] == pytest.approx(kp, rel=1e-12)
    assert d["max_abs_u"] == max(float(np.abs(x).max()) for x in u)
    assert d["nonfinite"] == 0


def test_nonfinite_is_reported():
    with arkamd.Solver(8, 8, 8) as s:
        s.init_tgv()
        u = s.download("u1nCon")
        u[3, 4, 5] = np.nan
        s.upload("u1nCon", u)
        with pytest.raises(arkamd.ArkError) as e:
            s.diagnostics()
        assert e.value.status == arkamd.ARK_ERR_NONFINITE
        assert s.diagnostics(allow_nonfinite=True)["nonfinite"] == 1


def test_errors_are_reported():
    with arkamd.Solver(6, 5, 4) as s:
        with pytest.raises(ValueError):
            s.upload("u11", np.zeros((6, 7, 8)))
        cond = np.zeros(s.shape("condition"))
        cond[2, 2, 2] = 1.0
        s.upload("condition", cond)
        # solid cells run on the phase path; the fused-only entry refuses them
        with pytest.raises(arkamd.ArkError) as e:
            s.advance(1)
        assert e.value.status == arkamd.ARK_ERR_UNSUPPORTED
    with pytest.raises(arkamd.ArkError) as e:
        arkamd.Solver(0, 4, 4, dx=(1.0, 1.0, 1.0))
    assert e.value.status == arkamd.ARK_ERR_INVALID


# ---- corrected scheme (gq from Q on every axis) ------------------------------
def test_corrected_scheme_phases_vs_oracle():
    st = random_fluid_state((13, 9, 7), 21)
    st.scheme = 1
    for phase in PHASES:
        want = st.copy()
        want.phase(phase)
        with arkamd.Solver(*st.N, dx=st.dx, CFL=st.CFL, VIS=st.VIS, sound=st.sound, ro0=st.ro0, t0=st.t0,
                           scheme=arkamd.SCHEME_CORRECTED) as s:
            s.upload_all(st.arrays)
            s.dt = st.dt
            if phase == "TimeStepSize":
                assert s.time_step_size() == want.dt
            else:
                s.phase(phase)
            assert_state_equal(s, want, what=f"corrected {phase}")


@pytest.mark.parametrize("N", [(24, 20, 16), (40, 33, 9)])
def test_corrected_scheme_fused_vs_oracle(N):
    ref = OracleState.tgv(*N, scheme=1)
    ref.step(6)
    with arkamd.Solver(*N, scheme=arkamd.SCHEME_CORRECTED) as s:
        s.init_tgv()
        s.step(6)
        assert_state_equal(s, ref, what=f"corrected fused {N}")


def test_reference_loop_by_hand_equals_step_phases():
    """The reference's loop written out through the C-ABI (source/main.cpp:
    82-91: TimeStepSize, nStep++, TIME += dt/2, the phases, TIME += dt/2)
    gives step_phases' arrays, TIME and nStep bit for bit."""
    with arkamd.Solver(16, 12, 10) as a, arkamd.Solver(16, 12, 10) as b:
        a.init_tgv()
        b.init_tgv()
        for _ in range(3):
            a.time_step_size()
            a.next_step()
            a.half_time()
            for ph in ("Phase1", "StressTensor", "UseForces", "Phase2", "Phase1", "UseForces"):
                a.phase(ph)
            a.half_time()
        b.step_phases(3)
        ga, gb = a.download_all(), b.download_all()
        bad = [describe_mismatch(n, ga[n], gb[n]) for n in FIELDS if not bits_equal(ga[n], gb[n])]
        assert not bad, "\n".join(bad)
        da, db = a.diagnostics(), b.diagnostics()
        assert da["time"] == db["time"] and da["nstep"] == db["nstep"] == 3 and da["dt"] == db["dt"]
