"""GPU: the BoundaryConditions and halo-plane entry points against the
reference's BoundaryConditions class (golden vectors) and its halo packing
loops (oracle restatement)."""
import numpy as np
import pytest

import arkamd
import ark_oracle as orc
from conftest import GOLDEN, bits_equal

pytestmark = pytest.mark.gpu

BC = np.load(f"{GOLDEN}/bc.npz")
CELLS = ["u1nCon", "u2nCon", "u3nCon", "ronCon", "tnCon"]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("periodic", [0, 1])
def test_boundary_conditions_vs_reference(axis, periodic):
    N1, N2, N3 = (int(v) for v in BC["N"])
    with arkamd.Solver(N1, N2, N3) as s:
        for name, a in zip(CELLS, BC["input"]):
            s.upload(name, a)
        s.boundary_conditions(axis, CELLS, bool(periodic), float(BC["t0"]))
        for q, name in enumerate(CELLS):
            assert bits_equal(s.download(name), BC[f"out_{axis}_{periodic}"][q]), (axis, periodic, name)


@pytest.mark.parametrize("N", [(7, 5, 6), (16, 12, 10)])
def test_planes_roundtrip(N):
    rng = np.random.default_rng(11)
    with arkamd.Solver(*N) as s:
        data = {}
        for name in ("u1Con", "ronCon", "u12", "p3"):
            data[name] = rng.normal(size=s.shape(name))
            s.upload(name, data[name])
        for name, a in data.items():
            ext = 1 if name.endswith("Con") else 2
            for axis in range(3):
                Nd = N[axis]
                for q in (0, 1, Nd, Nd + ext):
                    got = s.get_plane(name, axis, q)
                    assert bits_equal(got, orc.pack_plane(a, N, axis, q)), (name, axis, q)
        # set a plane, read the whole array back: only that plane's interior changed
        a = data["u1Con"]
        new = rng.normal(size=s.plane_shape(1))
        s.set_plane("u1Con", 1, N[1] + 1, new)
        b = a.copy()
        b[1:N[2] + 1, N[1] + 1, 1:N[0] + 1] = new
        assert bits_equal(s.download("u1Con"), b)
        with pytest.raises(arkamd.ArkError):
            s.get_plane("u1Con", 0, N[0] + 2)          # beyond a cell array
        with pytest.raises(arkamd.ArkError):
            s.set_plane("u1Con", 2, N[2] + 2, np.zeros(s.plane_shape(2)))


def test_planes_as_a_periodic_halo_exchange():
    """Packing the boundary planes and unpacking them on the opposite side (what
    the reference's MPI exchange does with one rank and periods=1) equals the
    library's own periodic fill: the next Phase2 sees the same ghosts."""
    N = (10, 8, 6)
    with arkamd.Solver(*N) as a, arkamd.Solver(*N) as b:
        a.init_tgv()
        b.init_tgv()
        for s in (a, b):
            s.time_step_size()
            s.phase1()
            s.stress_tensor()
            s.use_forces()
        for name in ("u1nCon", "u2nCon", "u3nCon", "ronCon", "tnCon"):
            for axis in range(3):
                Nd = N[axis]
                b.set_plane(name, axis, 0, b.get_plane(name, axis, Nd))
                b.set_plane(name, axis, Nd + 1, b.get_plane(name, axis, 1))
        a.phase2()
        b.phase2()
        for name in ("u11", "p2", "u33", "t1"):
            assert bits_equal(a.download(name), b.download(name)), name
