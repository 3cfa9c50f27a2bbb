"""GPU: the learned-closure pipeline at a small size: forced DNS at 64^3
generates filtered training pairs (width 2), the network trains on them, and
coarse 32^3 runs from the filtered initial field run with and without the
closure."""
import numpy as np
import pytest

import arkamd
from arkamd import closure as C
import turb_oracle as T

pytestmark = pytest.mark.gpu

VIS, EPS, SEED = 1.0 / 800.0, 0.1, 21


@pytest.fixture(scope="module")
def pipeline():
    d = C.dns(64, steps=120, snapshots=3, width=2, vis=VIS, eps=EPS, seed=SEED, series_every=40)
    tr = C.train(d["X"], d["T"], steps=400, batch=8192, lr=2e-3, seed=1)
    yield d, tr
    tr["mlp"].close()


def test_dns_data(pipeline):
    d, _ = pipeline
    assert d["nonfinite"] == 0
    assert d["X"].shape == (3 * 32 ** 3, 10) and d["T"].shape == (3 * 32 ** 3, 6)
    assert np.all(d["X"][:, 9] == np.float32(2 * d["dx"]))
    # the subgrid stress tensor of a box filter is positive semi-definite on
    # its diagonal (a variance)
    assert (d["T"][:, :3] >= -1e-7).all()
    ke = [r["ke"] for r in d["series"]]
    assert all(np.isfinite(ke)) and all(r["eps"] > 0 for r in d["series"])


def test_init_velocity_keeps_the_field(pipeline):
    d, _ = pipeline
    ub0 = d["ub0"]
    with arkamd.Solver(32, 32, 32, VIS=VIS, precision=arkamd.PREC_MIXED) as s:
        s.init_velocity(ub0)
        got = [s.download(n)[1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")]
        assert max(np.abs(g - u).max() for g, u in zip(got, ub0)) < 1e-12
        assert np.abs(s.spectrum(17) - T.spectrum(list(ub0), 17)).max() < 1e-14


def test_training_learns_the_stresses(pipeline):
    _, tr = pipeline
    assert tr["loss_end"] < 0.5 * tr["loss0"], (tr["loss0"], tr["loss_end"])


def test_les_with_and_without_closure(pipeline):
    d, tr = pipeline
    plain = C.les(d["ub0"], d["t_end"], VIS, EPS, SEED)
    model = C.les(d["ub0"], d["t_end"], VIS, EPS, SEED, mlp=tr["mlp"], delta=2 * d["dx"])
    assert plain["nonfinite"] == 0 and model["nonfinite"] == 0
    assert model["time"] >= d["t_end"] and plain["time"] >= d["t_end"]
    assert not np.array_equal(plain["E"], model["E"])          # the closure acts
    Eref = C.spectrum_of(d["ub_end"], VIS)
    for r in (plain, model):
        assert np.isfinite(C.spectrum_error(r["E"], Eref, 8))


def _shell_energy(du):
    """Energy per shell round(|m|) of a (3, N, N, N) periodic field."""
    N = du.shape[-1]
    m = np.fft.fftfreq(N, 1.0 / N)
    kk = np.sqrt(m[:, None, None] ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2)
    shell = np.rint(kk).astype(int)
    e = sum(np.abs(np.fft.fftn(c)) ** 2 for c in du)
    return np.bincount(shell.ravel(), weights=e.ravel())


def test_scale_selective_closure_acts_on_high_shells_only(pipeline):
    """With a cutoff, the closure's acceleration keeps only the shells >= kcut:
    over one step from the same state, the velocity change it causes has (to
    second order in dt) no energy below kcut, while the whole closure's does,
    and above kcut the two agree."""
    d, tr = pipeline
    kc = 6

    def one_step(closure, kcut=0):
        with arkamd.Solver(32, 32, 32, VIS=VIS, precision=arkamd.PREC_MIXED) as s:
            s.init_velocity(d["ub0"])
            if closure:
                s.set_closure(tr["mlp"], 2 * d["dx"])
                if kcut:
                    s.set_closure_cutoff(kcut)
            s.step(1)
            return np.stack([s.download(n)[1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")])

    base = one_step(False)
    e_full = _shell_energy(one_step(True) - base)
    e_cut = _shell_energy(one_step(True, kc) - base)
    assert e_cut[:kc].sum() < 1e-3 * e_cut[kc:].sum(), (e_cut[:kc].sum(), e_cut[kc:].sum())
    assert e_full[:kc].sum() > 100 * e_cut[:kc].sum()
    assert abs(e_cut[kc:].sum() / e_full[kc:].sum() - 1) < 0.05


def test_a_posteriori_closure_beats_plain_coarse_run():
    """The kept closure at the experiment's own configuration: a 512^3 forced
    run at nu = 1/360 (Re_lambda about 100), 2000 steps, eight 8-step
    correction windows; the network trained on the coarse solver's own lag
    behind the filtered fine run (closure.correction_pairs, 20000 Adam steps
    of 65536), its amplitude calibrated on the windows (closure.calibrate).
    The 128^3 run must reproduce the filtered 512^3 spectrum within 5% on
    shells 1..32 at the final time, which the calibration does not see, and
    beat the plain coarse run (the fallback) and the network as trained.
    Measured on seeds 53 / 59 / 61: plain 13.2 / 16.7 / 16.0 %, as trained
    5.1 / 5.6 / 6.4 %, calibrated 3.7 / 4.5 / 4.7 % (alpha 1.75 / 1.25 / 1.5)."""
    vis, eps, seed = 1.0 / 360.0, 0.1, 53
    d = C.correction_pairs(512, 2000, 8, 4, vis, eps, seed, tau=0.01, horizon=8)
    assert d["nonfinite"] == 0 and len(d["windows"]) == 8
    assert d["X"].shape == (8 * 128 ** 3, 10) and d["T"].shape == (8 * 128 ** 3, 6)
    tr = C.train(d["X"], d["T"], steps=20000, batch=65536, lr=1e-3, seed=1)
    delta = 4 * d["dx"]
    try:
        Eref = C.spectrum_of(d["ub_end"], vis)
        plain = C.les(d["ub0"], d["t_end"], vis, eps, seed, tau=0.01)
        raw = C.les(d["ub0"], d["t_end"], vis, eps, seed, mlp=tr["mlp"], delta=delta, tau=0.01)
        cal = C.calibrate(d, tr, vis, eps, seed, delta, tau=0.01)
        model = C.les(d["ub0"], d["t_end"], vis, eps, seed, mlp=tr["mlp"], delta=delta, tau=0.01)
    finally:
        tr["mlp"].close()
    assert plain["nonfinite"] == 0 and raw["nonfinite"] == 0 and model["nonfinite"] == 0
    e_plain = C.spectrum_error(plain["E"], Eref, 32)
    e_raw = C.spectrum_error(raw["E"], Eref, 32)
    e_model = C.spectrum_error(model["E"], Eref, 32)
    print("spectrum error, shells 1..32: plain %.4f, closure as trained %.4f, calibrated (alpha %.2f) %.4f; "
          "calibration %s" % (e_plain, e_raw, cal["alpha"], e_model, cal["candidates"]))
    assert e_raw < 0.5 * e_plain, (e_plain, e_raw)
    assert e_model < 0.05 and e_model < e_raw, (e_plain, e_raw, e_model, cal)
    # the closure is cheap beside the fine run it replaces
    assert model["wall"] * 20 < d["wall"], (model["wall"], d["wall"])
