"""CPU: the oracle's BoundaryConditions and halo-plane restatements against the
golden vectors of the reference's own BoundaryConditions class
(tests/golden/bc.npz, written by tests/golden/make_fixtures.py from
source/BoundaryConditions.cpp and source/Arr3d.cpp)."""
import numpy as np
import pytest

import ark_oracle as orc
from conftest import GOLDEN, bits_equal

BC = np.load(f"{GOLDEN}/bc.npz")


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("periodic", [0, 1])
def test_boundary_conditions_match_reference(axis, periodic):
    arrs = [a.copy() for a in BC["input"]]
    orc.boundary_conditions(arrs, axis, bool(periodic), float(BC["t0"]))
    for q in range(5):
        assert bits_equal(arrs[q], BC[f"out_{axis}_{periodic}"][q]), (axis, periodic, q)


def test_boundary_fixture_pins_the_quirks():
    """The fixture would catch the plausible mistakes: along x1 the period flag
    is ignored (zero-gradient either way), and the periodic top plane along x3
    takes t from ro(., ., 1), not from t."""
    inp = BC["input"]
    assert bits_equal(BC["out_0_0"], BC["out_0_1"])
    out = BC["out_2_1"]
    N1, N2, N3 = (int(v) for v in BC["N"])
    assert bits_equal(out[4][N3 + 1, 1:N2 + 1, 1:N1 + 1], inp[3][1, 1:N2 + 1, 1:N1 + 1])
    assert not np.array_equal(out[4][N3 + 1, 1:N2 + 1, 1:N1 + 1], inp[4][1, 1:N2 + 1, 1:N1 + 1])
    # edges and corners are never written
    for axis in range(3):
        for per in (0, 1):
            o = BC[f"out_{axis}_{per}"]
            assert bits_equal(o[:, 0, 0, :], inp[:, 0, 0, :]) and bits_equal(o[:, -1, -1, :], inp[:, -1, -1, :])


def test_pack_plane_follows_the_reference_loop():
    """pack_plane against the packing loop itself (source/main.cpp:619-630):
    buf.elem(j, k) = u(1, j + 1, k + 1), Arr2d index k*(n2-1) + j."""
    N = (5, 4, 3)
    a = np.random.default_rng(3).normal(size=(N[2] + 2, N[1] + 2, N[0] + 2))
    buf = np.empty(N[1] * N[2])
    for j in range(N[1]):
        for k in range(N[2]):
            buf[k * N[1] + j] = a[k + 1, j + 1, 1]
    assert bits_equal(orc.pack_plane(a, N, 0, 1).ravel(), buf)
    buf = np.empty(N[0] * N[2])
    for i in range(N[0]):
        for k in range(N[2]):
            buf[k * N[0] + i] = a[k + 1, 2, i + 1]
    assert bits_equal(orc.pack_plane(a, N, 1, 2).ravel(), buf)
    buf = np.empty(N[0] * N[1])
    for i in range(N[0]):
        for j in range(N[1]):
            buf[j * N[0] + i] = a[3, j + 1, i + 1]
    assert bits_equal(orc.pack_plane(a, N, 2, 3).ravel(), buf)
