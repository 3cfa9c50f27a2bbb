"""Test infrastructure only: numpy restatements of the homogeneous-turbulence
functions of the library (seeded initial field, forcing modes and field,
spectrum, box filter, filtered gradients, dissipation).  They have no counterpart in the
reference (a channel solver); each is pinned by the closed forms and
published known-answer vectors its tests check, and serves as the checker
of the HIP kernels.  Nothing in the product imports this module.

Arrays are numpy (N3, N2, N1) (x fastest), interior cells only.
"""
from __future__ import annotations

import numpy as np

MASK32 = np.uint64(0xFFFFFFFF)


# ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) --------------------------
def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 on uint32 counter words, scalar key."""
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = 0x9E3779B9, 0xBB67AE85
    c = [np.asarray(x, dtype=np.uint64) & MASK32 for x in (c0, c1, c2, c3)]
    k0, k1 = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def gauss4(seed: int, c0, c1, c2, c3):
    """Four standard normals per counter (Box-Muller on (x+1)/2^32)."""
    r = philox4x32_10(c0, c1, c2, c3, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    out = []
    for h in range(2):
        u1 = (r[2 * h].astype(np.float64) + 1.0) * (1.0 / 4294967296.0)
        u2 = (r[2 * h + 1].astype(np.float64) + 1.0) * (1.0 / 4294967296.0)
        rad = np.sqrt(-2.0 * np.log(u1))
        ang = 2.0 * np.pi * u2
        out += [rad * np.cos(ang), rad * np.sin(ang)]
    return out


# ---- spectral helpers -------------------------------------------------------------
def wavevectors(N, L):
    """Physical wavevector components on the rfftn grid (axes 0,1,2 = z,y,x)."""
    N1, N2, N3 = N
    m1 = np.arange(N1 // 2 + 1)
    m2 = np.fft.fftfreq(N2, 1.0 / N2)
    m3 = np.fft.fftfreq(N3, 1.0 / N3)
    K3, K2, K1 = np.meshgrid(m3, m2, m1, indexing="ij")
    return (2 * np.pi / L[0] * K1, 2 * np.pi / L[1] * K2, 2 * np.pi / L[2] * K3), (K1, K2, K3)


def init_turbulence(N, L, seed: int, k0: float, urms: float):
    """The library's initial field (ark_init_turbulence): returns the cell
    velocities u[3] and, per face orientation d, the face velocities
    faces[d][3], all (N3, N2, N1) interior arrays (cell/face index 1..N)."""
    N1, N2, N3 = N
    n = N1 * N2 * N3
    q = np.arange(n, dtype=np.uint64)
    g = gauss4(seed, q & MASK32, q >> np.uint64(32), np.uint64(0x7A3C51), np.uint64(0))
    noise = [g[d].reshape(N3, N2, N1) for d in range(3)]
    U = [np.fft.rfftn(a) for a in noise]
    (k1, k2, k3), (m1, m2, m3) = wavevectors(N, L)
    kk = k1 * k1 + k2 * k2 + k3 * k3
    nyq = (2 * m1 == N1) | (2 * np.abs(m2) == N2) | (2 * np.abs(m3) == N3)
    km = np.sqrt(kk)
    amp = km * np.exp(-(km / k0) ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        dot = (k1 * U[0] + k2 * U[1] + k3 * U[2]) / kk
    kv = (k1, k2, k3)
    for d in range(3):
        U[d] = np.where((kk == 0) | nyq, 0.0, amp * (U[d] - dot * kv[d]))
    w = np.where((m1 == 0) | (2 * m1 == N1), 1.0, 2.0)
    e = sum((w * np.abs(x) ** 2).sum() for x in U)
    scale = urms / np.sqrt(e / 3.0)
    cells = [np.fft.irfftn(x, s=(N3, N2, N1)) * n * scale for x in U]
    faces = []
    Ns = (N1, N2, N3)
    ms = (m1, m2, m3)
    for d in range(3):
        ph = np.exp(-0.5j * 2 * np.pi * ms[d] / Ns[d])
        faces.append([np.fft.irfftn(x * ph, s=(N3, N2, N1)) * n * scale for x in U])
    return cells, faces


def spectrum(u, nbins: int):
    """ark_spectrum: shell sums of 1/2 |u_hat|^2 / N^6 over round(|m|)."""
    N3, N2, N1 = u[0].shape
    n = N1 * N2 * N3
    U = [np.fft.rfftn(a) for a in u]
    _, (m1, m2, m3) = wavevectors((N1, N2, N3), (1.0, 1.0, 1.0))
    b = np.floor(np.sqrt(m1 ** 2 + m2 ** 2 + m3 ** 2) + 0.5).astype(int)
    w = np.where((m1 == 0) | (2 * m1 == N1), 1.0, 2.0)
    e = w * 0.5 * sum(np.abs(x) ** 2 for x in U) / (n * n)
    return np.bincount(b.ravel(), weights=e.ravel(), minlength=nbins)[:nbins]


def box_filter(u, w: int):
    """ark_filter_box: block means of u_i and of u_i u_j over w^3 cells."""
    N3, N2, N1 = u[0].shape
    def blk(a):
        return a.reshape(N3 // w, w, N2 // w, w, N1 // w, w).mean(axis=(1, 3, 5))
    ub = [blk(a) for a in u]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = [blk(u[i] * u[j]) - ub[i] * ub[j] for i, j in pairs]
    return np.array(ub), np.array(tau)


def coarse_gradient(ub, h):
    """Central differences of the filtered velocity (component 3i + j)."""
    out = []
    for i in range(3):
        for j, ax in enumerate((2, 1, 0)):          # x, y, z are numpy axes 2, 1, 0
            out.append((np.roll(ub[i], -1, axis=ax) - np.roll(ub[i], 1, axis=ax)) / (2.0 * h[j]))
    return np.array(out)


# ---- forcing (ark_set_forcing) ---------------------------------------------------
def forced_modes():
    """The forced integer wavevectors m: every 1 <= |m|^2 <= 4 whose first
    non-zero component is positive (the conjugates complete the real field),
    in lexicographic order of (m_x, m_y, m_z)."""
    out = []
    for a in range(-2, 3):
        for b in range(-2, 3):
            for d in range(-2, 3):
                first = a if a != 0 else (b if b != 0 else d)
                if 1 <= a * a + b * b + d * d <= 4 and first > 0:
                    out.append((a, b, d))
    return out


def forcing_table(seed: int, counter: int, eps: float, delta: float, L):
    """The complex amplitudes a_n (16 x 3) of one forcing interval: per mode
    n, eight normals from Philox counters (n, counter, 0xf0c3, 0 / 1); the
    first three are the real part, the next three the imaginary part; each is
    projected normal to k_n, and the pair scaled together to the norm
    A = sqrt(eps / (16 delta)), so that the field injects eps on average over
    an interval delta.  `counter` is the step number (white forcing, delta =
    dt) or the index of the correlation interval."""
    modes = forced_modes()
    table = np.zeros((len(modes), 3), dtype=np.complex128)
    for n, m in enumerate(modes):
        g = gauss4(seed, n, counter, 0xF0C3, 0) + gauss4(seed, n, counter, 0xF0C3, 1)
        g = [float(x) for x in g]
        k = np.array([2 * np.pi / L[d] * m[d] for d in range(3)])
        re, im = np.array(g[0:3]), np.array(g[3:6])
        re = re - (k @ re) / (k @ k) * k
        im = im - (k @ im) / (k @ k) * k
        nrm = float(re @ re + im @ im)
        A = np.sqrt(eps / (delta * 16.0)) / np.sqrt(nrm)
        table[n] = A * re + 1j * A * im
    return table


def force_field(table, N, L):
    """F_d(x) = sum_n 2 Re(a_nd e^{i k_n . x}) at the cell centres
    x_d = (i - 1/2) L_d / N_d, i = 1..N_d; arrays (3, N3, N2, N1)."""
    N1, N2, N3 = N
    x = (np.arange(N1) + 0.5) * L[0] / N1
    y = (np.arange(N2) + 0.5) * L[1] / N2
    z = (np.arange(N3) + 0.5) * L[2] / N3
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    F = np.zeros((3,) + X.shape)
    for n, m in enumerate(forced_modes()):
        ph = np.exp(1j * (2 * np.pi * m[0] / L[0] * X + 2 * np.pi * m[1] / L[1] * Y + 2 * np.pi * m[2] / L[2] * Z))
        for d in range(3):
            F[d] += 2.0 * np.real(table[n, d] * ph)
    return F


def dissipation(u, dx, nu):
    """ark_dissipation: nu <sum_ij (d_j u_i)^2> with central differences."""
    acc = 0.0
    for i in range(3):
        for j, ax in enumerate((2, 1, 0)):
            d = (np.roll(u[i], -1, axis=ax) - np.roll(u[i], 1, axis=ax)) / (2.0 * dx[j])
            acc += (d * d).mean()
    return nu * acc
