"""Host-side closure logic (no GPU): the stress whose closure force is a given
acceleration, the Taylor-scale Reynolds number, and the amplitude
calibration's choice (with the coarse runs stubbed)."""
import numpy as np
import pytest

from arkamd.closure import stress_from_force, taylor_reynolds

PAIRS = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}
AXIS = {0: 2, 1: 1, 2: 0}   # component x, y, z -> array axis of (N3, N2, N1)


def closure_force(tau, dx):
    """-sum_j (tau_ij(+1) - tau_ij(-1)) / (2 dx) with periodic neighbours: the
    acceleration the library's closure kernel forms from a stress field."""
    out = np.zeros((3,) + tau.shape[1:])
    for i in range(3):
        for j in range(3):
            a = tau[PAIRS[(i, j)]]
            out[i] -= (np.roll(a, -1, axis=AXIS[j]) - np.roll(a, 1, axis=AXIS[j])) / (2 * dx)
    return out


@pytest.mark.parametrize("N", [(12, 10, 8), (16, 16, 16), (9, 7, 6)])
def test_stress_reproduces_a_representable_force(N):
    """Any force the closure operator can produce is reproduced to rounding."""
    rng = np.random.default_rng(3)
    dx = 0.27
    F = closure_force(rng.standard_normal((6,) + N), dx)
    T = stress_from_force(F, dx)
    assert np.abs(closure_force(T, dx) - F).max() < 1e-12 * np.abs(F).max()


def test_stress_is_the_projection_of_any_force():
    """For an arbitrary force the result is its part in the operator's range:
    applying the inversion to that part changes nothing, and what is left over
    is orthogonal to every representable force."""
    rng = np.random.default_rng(4)
    N, dx = (10, 12, 8), 0.3
    F = rng.standard_normal((3,) + N)
    Fp = closure_force(stress_from_force(F, dx), dx)
    assert np.abs(closure_force(stress_from_force(Fp, dx), dx) - Fp).max() < 1e-12 * np.abs(F).max()
    G = closure_force(rng.standard_normal((6,) + N), dx)
    assert abs(np.sum((F - Fp) * G)) < 1e-10 * np.sqrt(np.sum(F ** 2) * np.sum(G ** 2))


def test_single_mode_closed_form():
    """F = (sin z, 0, 0) on a cube: the stress is tau_13 = cos(z) dx / sin(dx)
    (minus its central difference in z is sin z), every other component 0."""
    N, dx = 16, 2 * np.pi / 16
    z = (np.arange(N) + 0.5) * dx
    F = np.zeros((3, N, N, N))
    F[0] = np.sin(z)[:, None, None]
    T = stress_from_force(F, dx)
    want = np.cos(z)[:, None, None] * dx / np.sin(dx)
    assert np.allclose(T[4], want, atol=1e-12)
    for q in (0, 1, 2, 3, 5):
        assert np.abs(T[q]).max() < 1e-12


def test_taylor_reynolds_closed_form():
    """u' = 1 per component in a (2 pi)^3 box (KE = 1.5 V), nu = 1/360,
    eps = 0.15 * 360 / 100^2 -> Re_lambda = sqrt(15 / (nu eps)) = 100."""
    V = (2 * np.pi) ** 3
    nu = 1.0 / 360.0
    eps = 15.0 / (nu * 100.0 ** 2)
    assert taylor_reynolds(1.5 * V, eps, nu, V) == pytest.approx(100.0, rel=1e-12)


class _Net:
    """Stands in for arkamd.MLP: records the output normalisation it is given."""

    def __init__(self):
        self.norm = None

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        self.norm = (np.asarray(out_mean), np.asarray(out_std))


def _calibration_case(monkeypatch, err_of, nonfinite_alphas=()):
    """closure.calibrate on eight windows at t_n = 1..8 of a run ending at
    t = 9 (the last quarter holds t_n = 7 and 8), with the coarse run replaced
    by one whose spectrum at window q is off by err_of(alpha, q) on every
    shell.  Returns the result and the windows each run was asked for."""
    from arkamd import closure as C
    nc, nsh = 16, 9
    windows = [{"t_n": float(q), "E_ref": np.ones(nsh)} for q in range(1, 9)]
    d = {"windows": windows, "t_end": 9.0, "ub0": np.zeros((3, nc, nc, nc))}
    tr = {"mlp": _Net(), "norm": (np.zeros(10), np.ones(10), np.full(6, 0.5), np.full(6, 2.0))}
    asked = []

    def fake_les(ub0, times, vis, eps, seed, mlp=None, delta=0.0, tau=0.01):
        a = tr["alpha"]
        asked.append(list(times))
        return {"E_at": [np.full(nsh, 1.0 + err_of(a, q)) for q in range(1, 9)],
                "nonfinite": 1 if a in nonfinite_alphas else 0}

    monkeypatch.setattr(C, "les", fake_les)
    return C.calibrate(d, tr, 1.0, 0.1, 7, 0.5), tr, asked


def test_calibration_judges_the_last_quarter_only(monkeypatch):
    """Early windows favour alpha = 1, the last quarter alpha = 2: the
    calibration follows the last quarter, sets that amplitude on the network
    (mean and std of the outputs scaled by it), and every candidate run passes
    through all eight window times."""
    def err(a, q):
        return abs(a - 1.0) + 0.01 if q <= 6 else abs(a - 2.0) + 0.01
    res, tr, asked = _calibration_case(monkeypatch, err)
    assert res["alpha"] == 2.0 and tr["alpha"] == 2.0
    out_mean, out_std = tr["mlp"].norm
    assert np.allclose(out_mean, 1.0) and np.allclose(out_std, 4.0)
    assert all(t == [float(q) for q in range(1, 9)] for t in asked)
    assert len(asked) == len(res["candidates"]) == 7
    best = min(res["candidates"], key=lambda c: c["err_windows"])
    assert best["alpha"] == 2.0 and best["err_windows"] == pytest.approx(0.01)


def test_calibration_skips_runs_that_blow_up(monkeypatch):
    """A candidate whose run turned non-finite is never chosen, however small
    its error looks; with every candidate non-finite the network keeps alpha 1."""
    def err(a, q):
        return abs(a - 3.0) + 0.01
    res, tr, _ = _calibration_case(monkeypatch, err, nonfinite_alphas=(3.0, 2.5))
    assert res["alpha"] == 2.0
    res, tr, _ = _calibration_case(monkeypatch, err, nonfinite_alphas=(1.0, 1.25, 1.5, 1.75, 2.0, 2.5, 3.0))
    assert res["alpha"] == 1.0 and tr["alpha"] == 1.0
