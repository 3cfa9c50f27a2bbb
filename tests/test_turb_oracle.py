"""CPU: the turbulence oracle (oracle/turb_oracle.py) against published
known-answer vectors and closed forms."""
import numpy as np
import pytest

import turb_oracle as T


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of its authors (Random123 kat_vectors)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        got = T.philox4x32_10(*[np.uint64(c) for c in ctr], *key)
        assert tuple(int(x) for x in got) == want


def test_gauss_moments():
    q = np.arange(1 << 18, dtype=np.uint64)
    g = np.concatenate(T.gauss4(99, q, 0, 1, 2))
    assert abs(g.mean()) < 5e-3 and abs(g.var() - 1.0) < 5e-3
    assert abs((g ** 4).mean() - 3.0) < 3e-2          # Gaussian kurtosis


def test_single_mode_spectrum_and_parseval():
    N = 32
    x = (np.arange(N) + 0.5) * 2 * np.pi / N
    X = np.broadcast_to(x[None, None, :], (N, N, N))
    A, m = 0.7, 5
    u = [A * np.sin(m * X), np.zeros((N, N, N)), np.zeros((N, N, N))]
    E = T.spectrum(u, 30)
    assert abs(E[m] - A * A / 4) < 1e-14 and abs(E.sum() - E[m]) < 1e-14
    rng = np.random.default_rng(2)
    v = [rng.normal(size=(N, 24, 16)) for _ in range(3)]
    E = T.spectrum(v, 40)
    assert abs(E.sum() - 0.5 * sum((a * a).mean() for a in v)) < 1e-12


def test_init_turbulence_properties():
    N, L = (32, 32, 32), (2 * np.pi,) * 3
    cells, faces = T.init_turbulence(N, L, seed=7, k0=4.0, urms=1.0)
    # per-component rms
    assert abs(np.sqrt(sum((a * a).mean() for a in cells) / 3) - 1.0) < 1e-12
    # solenoidal: k . u_hat = 0 for every mode
    (k1, k2, k3), _ = T.wavevectors(N, L)
    U = [np.fft.rfftn(a) for a in cells]
    div = k1 * U[0] + k2 * U[1] + k3 * U[2]
    assert np.abs(div).max() < 1e-10 * max(np.abs(x).max() for x in U)
    # shell spectrum follows k^4 exp(-2 (k/k0)^2) within sampling noise
    E = T.spectrum(cells, 17)
    k = np.arange(17)
    ref = k ** 4.0 * np.exp(-2 * (k / 4.0) ** 2)
    r = E[3:12] / ref[3:12]
    assert r.max() / r.min() < 1.6
    # the faces are the same Fourier series half a cell down: compare the
    # x-face field with the cell field shifted by interpolation in Fourier space
    for d, ax in enumerate((2, 1, 0)):
        mid = 0.5 * (cells[0] + np.roll(cells[0], 1, axis=ax))   # second-order estimate of the face value
        err = np.abs(faces[d][0] - mid).max() / np.abs(cells[0]).max()
        assert err < 0.15


def test_face_shift_is_exact_for_a_single_mode():
    """One Fourier mode: the shifted inverse transform equals the function at
    the face positions (face q at (q-1) dx, cell q at (q-1/2) dx)."""
    N = 16
    xc = (np.arange(N) + 0.5) * 2 * np.pi / N
    xf = np.arange(N) * 2 * np.pi / N
    u = np.broadcast_to(np.cos(3 * xc)[None, None, :], (N, N, N))
    Uh = np.fft.rfftn(u)
    m1 = np.arange(N // 2 + 1)[None, None, :]
    f = np.fft.irfftn(Uh * np.exp(-0.5j * 2 * np.pi * m1 / N), s=(N, N, N))
    assert np.abs(f - np.cos(3 * xf)[None, None, :]).max() < 1e-13


def test_box_filter_transfer_function():
    """Mean of w samples of sin(m x): G = sin(m w dx / 2) / (w sin(m dx / 2)) times
    sin at the block centre."""
    N, w, m = 32, 4, 3
    dx = 2 * np.pi / N
    x = (np.arange(N) + 0.5) * dx
    u = [np.broadcast_to(np.sin(m * x)[None, None, :], (N, N, N)).copy(), np.ones((N, N, N)), np.zeros((N, N, N))]
    ub, tau = T.box_filter(u, w)
    G = np.sin(m * w * dx / 2) / (w * np.sin(m * dx / 2))
    xc = (np.arange(N // w) * w + w / 2) * dx
    assert np.abs(ub[0] - G * np.sin(m * xc)[None, None, :]).max() < 1e-14
    assert np.abs(tau[1]).max() < 1e-15 and np.abs(tau[3]).max() < 1e-15   # constant component
    blk = np.sin(m * x).reshape(-1, w)
    t11 = (blk ** 2).mean(axis=1) - blk.mean(axis=1) ** 2
    assert np.abs(tau[0] - t11[None, None, :]).max() < 1e-14


def test_dissipation_closed_form():
    N, nu, A, k = 24, 0.01, 0.3, 2
    dx = 2 * np.pi / N
    y = (np.arange(N) + 0.5) * dx
    u = [np.broadcast_to(A * np.sin(k * y)[None, :, None], (N, N, N)).copy(), np.zeros((N, N, N)),
         np.zeros((N, N, N))]
    want = nu * A * A * np.sin(k * dx) ** 2 / (2 * dx * dx)
    assert abs(T.dissipation(u, (dx, dx, dx), nu) - want) < 1e-14


def test_coarse_gradient_closed_form():
    """coarse_gradient's component 3i + j is the central difference of u_i
    along axis j: one single-mode field per component, each varying along a
    different axis, on a grid with a different spacing per axis, so that a
    swapped component order, axis map or spacing fails."""
    N3, N2, N1 = 8, 10, 12
    h = (2 * np.pi / N1, 2 * np.pi / N2, 2 * np.pi / N3)
    x = (np.arange(N1) + 0.5) * h[0]
    y = (np.arange(N2) + 0.5) * h[1]
    z = (np.arange(N3) + 0.5) * h[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ub = np.array([np.sin(2 * Y), 3 * np.sin(Z), 0.5 * np.sin(3 * X)])
    g = T.coarse_gradient(ub, h)
    want = np.zeros((9,) + X.shape)
    want[3 * 0 + 1] = np.cos(2 * Y) * np.sin(2 * h[1]) / h[1]          # d u_x / d y
    want[3 * 1 + 2] = 3 * np.cos(Z) * np.sin(h[2]) / h[2]              # d u_y / d z
    want[3 * 2 + 0] = 0.5 * np.cos(3 * X) * np.sin(3 * h[0]) / h[0]    # d u_z / d x
    assert g.shape == want.shape
    assert np.abs(g - want).max() < 1e-13


def test_forced_modes_closed_form():
    """32 integer vectors have 1 <= |m|^2 <= 4 (6 + 12 + 8 + 6 for |m|^2 =
    1, 2, 3, 4); one of each +-pair is forced."""
    m = T.forced_modes()
    assert len(m) == 16 and len(set(m)) == 16
    for v in m:
        assert 1 <= sum(c * c for c in v) <= 4
        assert tuple(-c for c in v) not in m
        assert next(c for c in v if c != 0) > 0
    assert sorted(m) == m


def test_forcing_table_and_field_closed_forms():
    """Each amplitude is normal to its wavevector and of norm
    sqrt(eps / (16 delta)); the field's grid mean square is then exactly
    sum_n 2 |a_n|^2 = 2 eps / delta (distinct wavevectors are orthogonal on
    the grid), its divergence is zero, and the table depends on seed and
    counter only."""
    L = (2 * np.pi, 2 * np.pi, 2 * np.pi)
    eps, delta = 0.1, 3e-3
    a = T.forcing_table(3, 1, eps, delta, L)
    for n, m in enumerate(T.forced_modes()):
        k = np.array(m, dtype=float)
        assert abs(k @ a[n]) < 1e-12 * np.linalg.norm(k)
        assert np.sqrt(np.sum(np.abs(a[n]) ** 2)) == pytest.approx(np.sqrt(eps / (16 * delta)), rel=1e-12)
    F = T.force_field(a, (12, 10, 8), L)
    assert np.sum(F ** 2, axis=0).mean() == pytest.approx(2 * eps / delta, rel=1e-12)
    assert np.array_equal(T.forcing_table(3, 1, eps, delta, L), a)
    assert not np.array_equal(T.forcing_table(3, 2, eps, delta, L), a)
    assert not np.array_equal(T.forcing_table(4, 1, eps, delta, L), a)


def test_force_field_single_mode():
    """One amplitude (1 + 2i) on mode n, component d: F_d = 2 (cos(k.x) - 2 sin(k.x))."""
    L = (2 * np.pi, 4 * np.pi, 2 * np.pi)
    N = (8, 6, 10)
    n = T.forced_modes().index((1, -1, 1))
    a = np.zeros((16, 3), dtype=complex)
    a[n, 1] = 1 + 2j
    F = T.force_field(a, N, L)
    x = (np.arange(N[0]) + 0.5) * L[0] / N[0]
    y = (np.arange(N[1]) + 0.5) * L[1] / N[1]
    z = (np.arange(N[2]) + 0.5) * L[2] / N[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    kx = 2 * np.pi / L[0] * X - 2 * np.pi / L[1] * Y + 2 * np.pi / L[2] * Z
    assert np.abs(F[1] - 2 * (np.cos(kx) - 2 * np.sin(kx))).max() < 1e-12
    assert np.abs(F[0]).max() == 0 and np.abs(F[2]).max() == 0
