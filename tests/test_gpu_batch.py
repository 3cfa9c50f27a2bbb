"""GPU: batched steps (one launch per step over many boxes, replayed from HIP
graphs, tile shape chosen by the autotuner).  Every box must evolve bit for
bit as its own standalone mixed-precision context does."""
import numpy as np
import pytest

import arkamd

pytestmark = pytest.mark.gpu

STATE = ["ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
         "u11", "u21", "u31", "p1", "t1", "u12", "u22", "u32", "p2", "t2", "u13", "u23", "u33", "p3", "t3"]
# uneven boxes and Reynolds numbers
BOXES = [((20, 16, 12), 1 / 400), ((32, 32, 32), 1 / 1600), ((24, 40, 16), 1 / 3200), ((48, 36, 30), 1 / 900)]


def make(boxes=BOXES, U0=1.0):
    out = []
    for N, vis in boxes:
        s = arkamd.Solver(*N, VIS=vis, precision=arkamd.PREC_MIXED)
        s.init_tgv(U0)
        out.append(s)
    return out


def assert_same(a, b, what):
    for sa, sb in zip(a, b):
        for n in STATE:
            x, y = sa.download(n), sb.download(n)
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), (what, sa.N, n)
        da, db = sa.diagnostics(), sb.diagnostics()
        assert da["dt"] == db["dt"] and da["nstep"] == db["nstep"] and da["time"] == db["time"], (what, sa.N)


def test_batch_equals_standalone():
    ref, bat = make(), make()
    for s in ref:
        s.step(7)
    with arkamd.Batch(bat) as b:
        b.step(7)
        cfg = b.config()
    assert cfg["tiles"] > 0
    assert_same(ref, bat, "batch vs standalone")


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("lz", [16, 64])
def test_batch_result_independent_of_tiling(shape, lz):
    ref, bat = make(BOXES[:2]), make(BOXES[:2])
    for s in ref:
        s.step(5)
    with arkamd.Batch(bat) as b:
        b.set_config(shape, lz)
        b.step(5)
    assert_same(ref, bat, f"shape {shape} lz {lz}")


def test_batch_graph_replays_with_eager_work_between():
    """Replays of the captured graphs, with the boxes' own eager work (a
    standalone step, which swaps its buffers and forces new tables, and
    downloads) between them, each checked against the standalone result."""
    ref, bat = make(BOXES[:3]), make(BOXES[:3])
    with arkamd.Batch(bat) as b:
        for k, (nb, solo) in enumerate([(3, False), (2, True), (1, False), (4, True), (2, False)]):
            b.step(nb)
            for s in ref:
                s.step(nb)
            assert_same(ref, bat, f"replay {k}")
            if solo:               # one box steps alone between batch calls
                bat[1].step(1)
                ref[1].step(1)
                bat[0].diagnostics()


def test_batch_autotune_leaves_state_unchanged():
    ref, bat = make(BOXES[:3]), make(BOXES[:3])
    for s in ref:
        s.step(2)
    with arkamd.Batch(bat) as b:
        b.step(1)
        ms = b.autotune(reps=2)
        assert len(ms) == 3 * len(arkamd.BATCH_SHAPES) and all(v > 0 for v in ms.values())
        best = min(ms, key=ms.get)
        cfg = b.config()
        assert (cfg["shape"], cfg["lz"]) == best
        b.step(1)
    assert_same(ref, bat, "after autotune")


def make_mixed_kinds(boxes, tau):
    """Corrected scheme; even boxes Taylor-Green, odd boxes seeded turbulence
    with random-phase forcing (per step for tau = 0, per interval otherwise)."""
    out = []
    for q, (N, vis) in enumerate(boxes):
        s = arkamd.Solver(*N, VIS=vis, precision=arkamd.PREC_MIXED, scheme=arkamd.SCHEME_CORRECTED)
        if q % 2:
            s.init_turbulence(100 + q, 2.0, 1.0)
            s.set_forcing(200 + q, 0.1, tau)
        else:
            s.init_tgv(1.0)
        out.append(s)
    return out


@pytest.mark.parametrize("tau", [0.0, 0.003])
def test_batch_with_forced_boxes_equals_standalone(tau):
    boxes = [((32, 32, 32), 1 / 500), ((32, 32, 32), 1 / 2000), ((48, 48, 48), 1 / 800), ((16, 32, 24), 1 / 3000)]
    ref, bat = make_mixed_kinds(boxes, tau), make_mixed_kinds(boxes, tau)
    for s in ref:
        s.step(9)
    with arkamd.Batch(bat) as b:
        b.step(4)
        b.step(5)
    assert_same(ref, bat, f"forced boxes, tau {tau}")
    for s in bat:
        assert s.diagnostics()["nonfinite"] == 0


def test_batch_rejects_bad_boxes():
    f64 = arkamd.Solver(16, 16, 16)
    m = arkamd.Solver(16, 16, 16, precision=arkamd.PREC_MIXED)
    with pytest.raises(arkamd.ArkError):
        arkamd.Batch([m, f64])
    with pytest.raises(arkamd.ArkError):
        arkamd.Batch([m, m])
    f64.close()
    m.close()


def test_batch_config4_scale_equals_standalone():
    """Boxes of the ensemble's larger edges (64, 96, 128), Taylor-Green and
    forced turbulence mixed, Re spread over [400, 3200], autotuned tiling:
    every box after 5 batched steps is bit for bit its standalone run."""
    boxes = [((64, 64, 64), 1 / 400), ((96, 96, 96), 1 / 1800), ((128, 128, 128), 1 / 3200),
             ((96, 96, 96), 1 / 700), ((64, 64, 64), 1 / 2600), ((128, 128, 128), 1 / 1200)]
    ref, bat = make_mixed_kinds(boxes, 0.01), make_mixed_kinds(boxes, 0.01)
    for s in ref:
        s.step(5)
    with arkamd.Batch(bat) as b:
        b.autotune(reps=1)
        b.step(2)
        b.step(3)
    assert_same(ref, bat, "config-4 scale")
    for s in bat:
        assert s.diagnostics()["nonfinite"] == 0
    for s in ref + bat:
        s.close()
