"""GPU: homogeneous turbulence (seeded initial field, forcing, spectrum, box
filter, dissipation) against the numpy oracle (oracle/turb_oracle.py) and
closed forms."""
import numpy as np
import pytest

import arkamd
import turb_oracle as T

pytestmark = pytest.mark.gpu


def cells_of(s):
    return [s.download(n)[1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")]


@pytest.mark.parametrize("N,prec", [((32, 32, 32), arkamd.PREC_FP64), ((24, 32, 16), arkamd.PREC_MIXED)])
def test_init_turbulence_vs_oracle(N, prec):
    with arkamd.Solver(*N, precision=prec) as s:
        s.init_turbulence(seed=11, k0=4.0, urms=1.0)
        L = tuple(d * n for d, n in zip(s.dx, N))
        cells, faces = T.init_turbulence(N, L, 11, 4.0, 1.0)
        for a, b in zip(cells_of(s), cells):
            assert np.abs(a - b).max() < 1e-12
        tol = 1e-12 if prec == arkamd.PREC_FP64 else 1e-6     # FP32 face storage
        for d in range(3):
            for m in range(3):
                f = s.download(["u11", "u21", "u31", "u12", "u22", "u32", "u13", "u23", "u33"][3 * d + m])
                got = f[1:N[2] + 1, 1:N[1] + 1, 1:N[0] + 1]
                assert np.abs(got - faces[d][m]).max() < tol, (d, m)
                # exactly periodic: face N+1 is face 1
                if d == 0:
                    assert np.array_equal(f[1:N[2] + 1, 1:N[1] + 1, N[0] + 1], f[1:N[2] + 1, 1:N[1] + 1, 1])
        assert np.all(s.download("ronCon") == 1.0)
        d = s.diagnostics()
        assert d["nstep"] == 0 and d["time"] == 0.0


def test_spectrum_dissipation_filter_vs_oracle():
    N = (32, 24, 16)
    with arkamd.Solver(*N, VIS=0.02) as s:
        s.init_turbulence(seed=5, k0=3.0, urms=0.8)
        u = cells_of(s)
        nb = 30
        assert np.abs(s.spectrum(nb) - T.spectrum(u, nb)).max() < 1e-14
        assert abs(s.spectrum(nb).sum() - 0.5 * sum((a * a).mean() for a in u)) < 1e-13
        assert abs(s.dissipation() - T.dissipation(u, s.dx, 0.02)) < 1e-13
        ub, tau, gr = s.filter_box(4)
        ub0, tau0 = T.box_filter(u, 4)
        gr0 = T.coarse_gradient(ub0, tuple(4 * d for d in s.dx))
        assert np.abs(ub - ub0).max() < 1e-14 and np.abs(tau - tau0).max() < 1e-14
        assert np.abs(gr - gr0).max() < 1e-12


@pytest.mark.parametrize("N", [(16, 32, 32), (32, 16, 32), (48, 40, 24)])
def test_turbulence_functions_are_reproducible(N):
    """Two contexts given the same seed hold the same field bit for bit, and
    the spectrum and dissipation repeat bit for bit (every sum over the grid
    is taken in a fixed order)."""
    out = []
    for _ in range(2):
        with arkamd.Solver(*N, VIS=0.01, precision=arkamd.PREC_MIXED) as s:
            s.init_turbulence(seed=21, k0=4.0, urms=1.0)
            out.append(([s.download(n) for n in ("u1nCon", "u2nCon", "u3nCon", "u11", "u22", "u33")],
                        s.spectrum(), s.dissipation()))
    (fa, ea, da), (fb, eb, db) = out
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    assert np.array_equal(ea.view(np.uint64), eb.view(np.uint64)) and da == db


def test_single_mode_spectrum_on_device():
    N = 32
    with arkamd.Solver(N, N, N) as s:
        a = np.zeros(s.shape("u1nCon"))
        x = (np.arange(N + 2) - 0.5) * s.dx[0]
        a[:] = 0.7 * np.sin(5 * x)[None, None, :]
        s.upload("u1nCon", a)
        for n in ("u2nCon", "u3nCon"):
            s.upload(n, np.zeros(s.shape(n)))
        E = s.spectrum(30)
        assert abs(E[5] - 0.49 / 4) < 1e-14 and abs(E.sum() - E[5]) < 1e-14


def rest(N, vis=1e-6):
    s = arkamd.Solver(*N, VIS=vis, precision=arkamd.PREC_MIXED)
    s.init_turbulence(seed=1, k0=4.0, urms=0.0)      # fluid at rest, uniform density
    return s


def test_forcing_injects_eps_dt_in_one_step():
    """From rest, one step of white forcing gives u = F dt, so the kinetic
    energy per unit mass is <|F|^2> dt^2 / 2 = eps dt (the grid mean of the
    forced modes' squares is exact).  Tolerance 1%: within the step the
    scheme's pressure coupling (Phase2 faces between the two half-step force
    applications) takes a few tenths of a percent."""
    eps = 0.1
    with rest((32, 32, 32)) as s:
        s.set_forcing(seed=3, eps=eps)
        s.step(1)
        d = s.diagnostics()
        ke = 0.5 * sum((a * a).mean() for a in cells_of(s))
        assert abs(ke / (eps * d["dt"]) - 1.0) < 1e-2, ke / (eps * d["dt"])


def _zero_network():
    """A closure network whose stresses are exactly zero (all parameters and
    the output offsets zero): attaching it switches a forced context to the
    force-field path without changing the force."""
    m = arkamd.MLP(seed=0)
    m.set_params(np.zeros(m.nparam, dtype=np.float32))
    m.set_normalization(np.zeros(10), np.ones(10), np.zeros(6), np.ones(6))
    return m


def test_forcing_field_matches_the_oracle():
    """From rest, one step of white forcing leaves u = F dt in every cell,
    with F the oracle's field of the step's mode table (counter: step 1, delta
    = dt): this pins the mode list, amplitudes, phases and signs of the
    forcing the fused step evaluates in place.  The force-field path (taken
    when a closure adds to the field; here one that adds zero) gives the same
    velocities to FP32 rounding.  Tolerances: 1e-2 against the oracle (within
    the step the scheme's pressure coupling moves u by a few tenths of a
    percent), 1e-5 between the two paths (FP32 rounding of F in two orders)."""
    N, eps = (32, 24, 16), 0.1
    L = (2 * np.pi,) * 3
    out = {}
    net = _zero_network()
    for path in ("in place", "field"):
        with rest(N) as s:
            s.set_forcing(seed=3, eps=eps)
            if path == "field":
                s.set_closure(net, 1.0)
            s.step(1)
            out[path] = (np.stack(cells_of(s)), s.diagnostics()["dt"])
    net.close()
    u, dt = out["in place"]
    F = T.force_field(T.forcing_table(3, 1, eps, dt, L), N, L)
    err = np.sqrt(np.sum((u - F * dt) ** 2) / np.sum((F * dt) ** 2))
    uf, dtf = out["field"]
    assert dtf == dt
    between = np.sqrt(np.sum((u - uf) ** 2) / np.sum(u ** 2))
    print(f"oracle {err:.2e}, paths {between:.2e}")
    assert err < 1e-2, err
    assert between < 1e-5, between


def test_forcing_with_correlation_time():
    """Amplitudes held over intervals of tau: from rest one step gives
    u = F dt with <|F|^2> = 2 eps / tau, so KE = eps dt^2 / tau; and the force
    depends on time only, not on the step count."""
    eps, tau = 0.1, 0.05
    with rest((32, 32, 32)) as s:
        s.set_forcing(seed=3, eps=eps, tau=tau)
        s.step(1)
        d = s.diagnostics()
        ke = 0.5 * sum((a * a).mean() for a in cells_of(s))
        assert abs(ke / (eps * d["dt"] ** 2 / tau) - 1.0) < 1e-2


def test_forcing_is_seeded_and_solenoidal():
    N = (32, 32, 32)
    out = []
    for seed in (3, 3, 4):
        with rest(N) as s:
            s.set_forcing(seed=seed, eps=0.1)
            s.step(3)
            out.append(cells_of(s))
    assert all(np.array_equal(a, b) for a, b in zip(out[0], out[1]))
    assert not np.array_equal(out[0][0], out[2][0])
    # only the forced shells 1 <= |m| <= 2 carry energy; the field is solenoidal
    E = T.spectrum(out[0], 17)
    assert E[1:3].sum() > 0.98 * E.sum()
    (k1, k2, k3), _ = T.wavevectors(N, (2 * np.pi,) * 3)
    U = [np.fft.rfftn(a) for a in out[0]]
    div = np.abs(k1 * U[0] + k2 * U[1] + k3 * U[2]).max()
    assert div < 2e-2 * max(np.abs(x).max() for x in U)


def test_forced_turbulence_runs():
    with arkamd.Solver(32, 32, 32, VIS=1 / 400, precision=arkamd.PREC_MIXED) as s:
        s.init_turbulence(seed=2, k0=4.0, urms=1.0)
        s.set_forcing(seed=9, eps=0.1)
        s.step(200)
        d = s.diagnostics()
        assert d["nonfinite"] == 0 and d["nstep"] == 200
        assert s.dissipation() > 0


def test_forced_turbulence_reaches_a_stationary_state():
    """Long-run regression of the forcing and the corrected scheme at 64^3
    (seed 31, eps = 0.1 injected, nu = 1/360, 24 time units, about 10 000
    steps): after the initial field's transient the kinetic energy per unit
    volume stays bounded and stops drifting (the means of the last two
    quarters differ by less than 20%; it fluctuates by about +-20% within a
    quarter), so the total dissipation balances the injection on average,
    and the resolved viscous part of it is a fraction of the injection (about
    half at this grid; the rest is the scheme's grid-scale dissipation)."""
    vis, eps, V = 1.0 / 360.0, 0.1, (2 * np.pi) ** 3
    with arkamd.Solver(64, 64, 64, VIS=vis, precision=arkamd.PREC_MIXED, scheme=arkamd.SCHEME_CORRECTED) as s:
        s.init_turbulence(31, 4.0, 1.0)
        s.set_forcing(32, eps, 0.01)
        rec = []
        while True:
            d = s.diagnostics()
            rec.append((d["time"], d["ke_physical"] / V, s.dissipation()))
            if d["time"] >= 24.0:
                break
            s.step(200)
    t = np.array([r[0] for r in rec])
    ke = np.array([r[1] for r in rec])
    er = np.array([r[2] for r in rec])
    a, b = (t >= 12) & (t < 18), t >= 18
    assert np.isfinite(ke).all() and np.isfinite(er).all()
    assert ke[0] == pytest.approx(1.5, rel=1e-6)        # u_rms 1 per component
    assert 0.3 < ke[a | b].min() and ke[a | b].max() < 1.0
    assert abs(ke[b].mean() - ke[a].mean()) < 0.2 * ke[a | b].mean()
    share = er[a | b].mean() / eps
    assert 0.3 < share < 1.0, share


@pytest.mark.parametrize("prec", [arkamd.PREC_FP64, arkamd.PREC_MIXED])
def test_init_state_and_filter_density(prec):
    """ark_init_state puts a given density on the cells (and, shifted half a
    cell, on the faces, whose pressures follow the equation of state); the
    velocities are those of ark_init_velocity; ark_filter_density returns the
    block means of the cell density."""
    N = 16
    with arkamd.Solver(N, N, N, precision=prec) as a, arkamd.Solver(N, N, N, precision=prec) as b:
        a.init_turbulence(5, 3.0, 1.0)
        u = np.stack([a.download(n)[1:-1, 1:-1, 1:-1] for n in ("u1nCon", "u2nCon", "u3nCon")])
        x = (np.arange(N) + 0.5) * 2 * np.pi / N
        rho = 1.0 + 1e-3 * (np.sin(x)[None, None, :] * np.cos(2 * x)[:, None, None] + 0.5 * np.cos(x)[None, :, None])
        b.init_state(u, rho)
        a.init_velocity(u)
        for n in ("u1nCon", "u22", "u13"):
            assert np.array_equal(a.download(n), b.download(n)), n
        got = b.download("ronCon")[1:-1, 1:-1, 1:-1]
        assert np.abs(got - rho).max() < 1e-15
        # x-faces at the cell's low side: rho(x - dx/2) = 1 + 1e-3 (sin(x - dx/2) cos 2z + 0.5 cos y)
        xf = np.arange(N) * 2 * np.pi / N
        want = 1.0 + 1e-3 * (np.sin(xf)[None, None, :] * np.cos(2 * x)[:, None, None] + 0.5 * np.cos(x)[None, :, None])
        p1 = b.download("p1")[1:-2, 1:-2, 1:-2]
        assert np.abs(p1 - 100.0 * (want - 1.0)).max() < (1e-6 if prec == arkamd.PREC_MIXED else 1e-12)
        if prec == arkamd.PREC_FP64:   # face densities derived from p, bit for bit
            ro1 = b.download("ro1")[1:-2, 1:-2, 1:-2]
            assert np.array_equal(ro1, 1.0 + p1 / (10.0 * 10.0))
        fb = b.filter_density(4)
        assert np.abs(fb - got.reshape(4, 4, 4, 4, 4, 4).mean(axis=(1, 3, 5))).max() < 1e-15
