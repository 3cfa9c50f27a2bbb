"""Context state across calls: non-finite detection in the fused steps,
TimeStepSize between fused steps, face uploads after fused steps, and
checkpoint/restart.  Every comparison is bit for bit (FP64 and mixed
precision alike: the same arithmetic on the same inputs).
"""
import numpy as np
import pytest

from conftest import bits_equal, describe_mismatch
import arkamd

pytestmark = pytest.mark.gpu

PRECS = [arkamd.PREC_FP64, arkamd.PREC_MIXED]


def assert_same(a: arkamd.Solver, b: arkamd.Solver, names=arkamd.PERSISTENT, what=""):
    bad = []
    for n in names:
        x, y = a.download(n), b.download(n)
        if not bits_equal(x, y):
            bad.append(describe_mismatch(n, x, y))
    assert not bad, what + "\n" + "\n".join(bad)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("where", ["u1nCon", "p2"])
def test_nonfinite_found_while_stepping(prec, where):
    """A NaN uploaded into one cell (or face) reaches the half level of the
    next step, whose CFL reduction counts it: the next synchronising call
    reports ARK_ERR_NONFINITE naming step 1, and stepping is refused until a
    new initial state.  (An uploaded FP64 pressure face makes the first step
    a phase step, whose count comes from the next step's TimeStepSize scan.)"""
    N = (20, 18, 16)
    with arkamd.Solver(*N, precision=prec, scheme=arkamd.SCHEME_CORRECTED) as s:
        s.init_tgv()
        a = s.download(where)
        a[5, 6, 7] = np.nan
        s.upload(where, a)
        try:
            s.step(2) if where.startswith("p") else s.advance(2)
        except arkamd.ArkError as e:    # allowed: the count may already be back
            assert e.status == arkamd.ARK_ERR_NONFINITE
        with pytest.raises(arkamd.ArkError) as e:
            s.synchronize()
        assert e.value.status == arkamd.ARK_ERR_NONFINITE
        assert "first seen in step 1" in str(e.value), str(e.value)
        with pytest.raises(arkamd.ArkError) as e:
            s.advance(1)
        assert e.value.status == arkamd.ARK_ERR_NONFINITE
        with pytest.raises(arkamd.ArkError) as e:
            s.step(1)
        assert e.value.status == arkamd.ARK_ERR_NONFINITE
        s.init_tgv()        # a new initial state clears it
        s.advance(3)
        s.synchronize()
        assert s.diagnostics()["nstep"] == 3


def test_finite_run_reports_nothing():
    with arkamd.Solver(24, 24, 24, precision=arkamd.PREC_MIXED) as s:
        s.init_tgv()
        for _ in range(5):
            s.step(3)
        s.synchronize()
        assert s.diagnostics()["nonfinite"] == 0


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_time_step_size_between_fused_steps(prec):
    """TimeStepSize after fused steps returns the next step's dt and leaves it
    pending, as calling TimeStepSize twice in the reference gives the same dt:
    advance(3); time_step_size(); advance(2) equals advance(5)."""
    N = (24, 20, 18)
    with arkamd.Solver(*N, precision=prec) as a, arkamd.Solver(*N, precision=prec) as b:
        a.init_tgv()
        b.init_tgv()
        a.advance(5)
        b.advance(3)
        dt1 = b.time_step_size()
        dt2 = b.time_step_size()
        assert dt1 == dt2
        b.advance(2)
        assert_same(a, b, what="advance(5) vs advance(3) + TimeStepSize + advance(2)")
        da, db = a.diagnostics(), b.diagnostics()
        assert da["dt"] == db["dt"] and da["time"] == db["time"] and da["nstep"] == db["nstep"] == 5


def test_time_step_size_between_mixed_steps():
    N = (24, 20, 18)
    with arkamd.Solver(*N, precision=arkamd.PREC_MIXED) as a, \
            arkamd.Solver(*N, precision=arkamd.PREC_MIXED) as b:
        a.init_tgv()
        b.init_tgv()
        a.step(5)
        b.step(3)
        b.time_step_size()
        b.step(2)
        assert_same(a, b, what="step(5) vs step(3) + TimeStepSize + step(2)")


def test_face_upload_after_advance():
    """advance(3) leaves the face densities derived from p; uploading p1 (its
    own values) must first store them, so the phase step that follows reads
    the reference's face densities: equal to advance(3), step(2)."""
    N = (20, 16, 14)
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        a.advance(3)
        b.advance(3)
        b.upload("p1", b.download("p1"))
        a.step(2)
        b.step(2)
        assert_same(a, b, names=arkamd.FIELDS, what="face upload after advance")


def test_plane_upload_after_advance():
    N = (20, 16, 14)
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        a.advance(3)
        b.advance(3)
        plane = b.get_plane("p2", 1, 5)
        b.set_plane("p2", 1, 5, plane)
        a.step(2)
        b.step(2)
        assert_same(a, b, names=arkamd.FIELDS, what="plane upload after advance")


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("how", ["advance", "step"])
def test_restart_is_bitwise(prec, how):
    """N steps equal N/2 steps, a checkpoint (the persistent state and the
    clock), a fresh context resumed from it, and N/2 more steps."""
    N, steps = (26, 22, 18), 12
    run = (lambda s, n: s.advance(n)) if how == "advance" else (lambda s, n: s.step(n))
    with arkamd.Solver(*N, precision=prec, scheme=arkamd.SCHEME_CORRECTED) as a:
        a.init_tgv()
        run(a, steps)
        want = a.download_all(arkamd.PERSISTENT)
        want_clock = a.clock()
    with arkamd.Solver(*N, precision=prec, scheme=arkamd.SCHEME_CORRECTED) as b:
        b.init_tgv()
        run(b, steps // 2)
        ck = b.checkpoint()
    with arkamd.Solver(*N, precision=prec, scheme=arkamd.SCHEME_CORRECTED) as c:
        c.resume(ck)
        assert c.clock() == ck["clock"]
        run(c, steps - steps // 2)
        got = c.download_all(arkamd.PERSISTENT)
        clock = c.clock()
    bad = [describe_mismatch(n, got[n], want[n]) for n in arkamd.PERSISTENT if not bits_equal(got[n], want[n])]
    assert not bad, "\n".join(bad)
    assert clock == want_clock


def test_restart_rejects_bad_dt():
    with arkamd.Solver(8, 8, 8) as s:
        with pytest.raises(arkamd.ArkError) as e:
            s.restart({"time": 0.0, "nstep": 0, "dt": 0.0, "dt_next": -1.0, "dt_next_known": 1})
        assert e.value.status == arkamd.ARK_ERR_INVALID


def test_channel_refused_in_mixed_precision():
    with arkamd.Solver(8, 8, 8, precision=arkamd.PREC_MIXED) as s:
        with pytest.raises(arkamd.ArkError) as e:
            s.init_channel(1.0)
        assert e.value.status == arkamd.ARK_ERR_UNSUPPORTED
