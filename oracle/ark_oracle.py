"""Test infrastructure only: ctypes front end of the oracle (``libark_oracle.so``).

The oracle is a CPU restatement of the reference step (``oracle/ark_oracle.c``);
tests use it as the checker for the HIP path.  Nothing in the product imports
this module.

Arrays use the reference's layout and extents (source/main.cpp:198-277,
source/Arr3d.cpp:19-21): numpy shape (N3+2, N2+2, N1+2) for cell fields and
(N3+3, N2+3, N1+3) for face and stress fields, C order, so x is fastest.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libark_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ark_ref_harness")

FIELDS = [
    "roCon", "u1Con", "u2Con", "u3Con", "tCon",
    "ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
    "f1", "f2", "f3",
    "condition",
    "ro1", "t1", "u11", "u21", "u31", "p1",
    "ro2", "t2", "u12", "u22", "u32", "p2",
    "ro3", "t3", "u13", "u23", "u33", "p3",
    "sigm11", "sigm21", "sigm31",
    "sigm12", "sigm22", "sigm32",
    "sigm13", "sigm23", "sigm33",
]
FACE_FIRST = FIELDS.index("ro1")
CELL_STATE = ["ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon"]
FACE_STATE = ["u11", "u21", "u31", "p1", "t1", "ro1",
              "u12", "u22", "u32", "p2", "t2", "ro2",
              "u13", "u23", "u33", "p3", "t3", "ro3"]


def is_face(name: str) -> bool:
    return FIELDS.index(name) >= FACE_FIRST


class _State(ctypes.Structure):
    _fields_ = [
        ("N1", ctypes.c_int), ("N2", ctypes.c_int), ("N3", ctypes.c_int),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double), ("dx3", ctypes.c_double),
        ("dt", ctypes.c_double), ("CFL", ctypes.c_double), ("VIS", ctypes.c_double),
        ("sound", ctypes.c_double), ("ro0", ctypes.c_double), ("t0", ctypes.c_double),
        ("f", ctypes.POINTER(ctypes.c_double) * len(FIELDS)),
        ("scheme", ctypes.c_int),
        ("boundary", ctypes.c_int),
    ]


_lib = None


def build() -> None:
    """Compile the oracle library (plain C, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "libark_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(_State)
        for fn in ("ark_oracle_time_step_size", "ark_oracle_phase1", "ark_oracle_stress_tensor",
                   "ark_oracle_use_forces", "ark_oracle_phase2"):
            getattr(L, fn).argtypes = [P]
            getattr(L, fn).restype = None
        L.ark_oracle_step.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.ark_oracle_step.restype = None
        L.ark_oracle_init_tgv.argtypes = [P, ctypes.c_double]
        L.ark_oracle_init_tgv.restype = None
        L.ark_oracle_init_channel.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ark_oracle_init_channel.restype = None
        L.ark_oracle_energy.argtypes = [P]
        L.ark_oracle_energy.restype = ctypes.c_double
        L.ark_oracle_mass.argtypes = [P]
        L.ark_oracle_mass.restype = ctypes.c_double
        _lib = L
    return _lib


class OracleState:
    """Full solver state held as numpy arrays in the reference layout."""

    def __init__(self, N1: int, N2: int, N3: int, dx=(1.0, 1.0, 1.0), dt: float = 0.0,
                 CFL: float = 0.3, VIS: float = 1.0 / 1600.0, sound: float = 10.0,
                 ro0: float = 1.0, t0: float = 1.0, scheme: int = 0, boundary: int = 0):
        self.N = (int(N1), int(N2), int(N3))
        self.scheme = int(scheme)
        self.boundary = int(boundary)
        self.dx = tuple(float(v) for v in dx)
        self.dt, self.CFL, self.VIS = float(dt), float(CFL), float(VIS)
        self.sound, self.ro0, self.t0 = float(sound), float(ro0), float(t0)
        self.arrays: Dict[str, np.ndarray] = {}
        for name in FIELDS:
            self.arrays[name] = np.zeros(self.shape(name), dtype=np.float64)

    # ---- layout -----------------------------------------------------------
    def shape(self, name: str):
        N1, N2, N3 = self.N
        e = 3 if is_face(name) else 2
        return (N3 + e, N2 + e, N1 + e)

    def __getitem__(self, name: str) -> np.ndarray:
        return self.arrays[name]

    def __setitem__(self, name: str, value) -> None:
        a = np.ascontiguousarray(value, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.shape(name)}")
        self.arrays[name] = a.copy()

    def copy(self) -> "OracleState":
        o = OracleState(*self.N, dx=self.dx, dt=self.dt, CFL=self.CFL, VIS=self.VIS,
                        sound=self.sound, ro0=self.ro0, t0=self.t0, scheme=self.scheme,
                        boundary=self.boundary)
        for k, v in self.arrays.items():
            o.arrays[k] = v.copy()
        return o

    def _struct(self) -> _State:
        s = _State()
        s.N1, s.N2, s.N3 = self.N
        s.dx1, s.dx2, s.dx3 = self.dx
        s.dt, s.CFL, s.VIS, s.sound, s.ro0, s.t0 = (self.dt, self.CFL, self.VIS,
                                                   self.sound, self.ro0, self.t0)
        s.scheme = self.scheme
        s.boundary = self.boundary
        for q, name in enumerate(FIELDS):
            a = self.arrays[name]
            assert a.flags.c_contiguous and a.dtype == np.float64
            s.f[q] = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return s

    def _call(self, fn: str) -> None:
        s = self._struct()
        getattr(lib(), fn)(ctypes.byref(s))
        self.dt = s.dt

    # ---- the reference's phase functions ------------------------------------
    def time_step_size(self) -> float:
        self._call("ark_oracle_time_step_size")
        return self.dt

    def phase1(self) -> None:
        self._call("ark_oracle_phase1")

    def stress_tensor(self) -> None:
        self._call("ark_oracle_stress_tensor")

    def use_forces(self) -> None:
        self._call("ark_oracle_use_forces")

    def phase2(self) -> None:
        self._call("ark_oracle_phase2")

    def phase(self, name: str) -> None:
        {"TimeStepSize": self.time_step_size, "Phase1": self.phase1,
         "StressTensor": self.stress_tensor, "UseForces": self.use_forces,
         "Phase2": self.phase2, "Step": lambda: self.step(1)}[name]()

    def step(self, nsteps: int = 1) -> np.ndarray:
        s = self._struct()
        dts = (ctypes.c_double * max(nsteps, 1))()
        lib().ark_oracle_step(ctypes.byref(s), nsteps, dts)
        self.dt = s.dt
        return np.array(dts[:nsteps])

    def energy_sum(self) -> float:
        s = self._struct()
        return lib().ark_oracle_energy(ctypes.byref(s))

    def mass(self) -> float:
        s = self._struct()
        return lib().ark_oracle_mass(ctypes.byref(s))

    # ---- constructors -------------------------------------------------------
    @classmethod
    def tgv(cls, N1: int, N2: Optional[int] = None, N3: Optional[int] = None,
            VIS: float = 1.0 / 1600.0, CFL: float = 0.3, sound: float = 10.0,
            U0: float = 1.0, scheme: int = 0) -> "OracleState":
        N2 = N1 if N2 is None else N2
        N3 = N1 if N3 is None else N3
        two_pi = 2.0 * np.pi
        dx = ((two_pi - 0.0) / N1, (two_pi - 0.0) / N2, (two_pi - 0.0) / N3)
        st = cls(N1, N2, N3, dx=dx, CFL=CFL, VIS=VIS, sound=sound, scheme=scheme)
        s = st._struct()
        lib().ark_oracle_init_tgv(ctypes.byref(s), U0)
        return st

    @classmethod
    def channel(cls, n1_g: int = 201, n2_g: int = 61, n3_g: int = 2, x1_e: float = 20.0, x2_n: float = 6.0,
                x3_t: float = 0.1, u10: float = 1.0, VIS: float = 1.0 / 1000.0, CFL: float = 0.2,
                sound: float = 10.0) -> "OracleState":
        """The reference's own problem (source/main.cpp:118-372; defaults are
        its Input() values): the channel with a backward-facing step, inflow
        at x = 0, outflow at x = x1_e, walls at y = 0 and y = x2_n."""
        st = cls(n1_g - 1, n2_g - 1, n3_g - 1, CFL=CFL, VIS=VIS, sound=sound, boundary=1)
        s = st._struct()
        lib().ark_oracle_init_channel(ctypes.byref(s), n1_g, n2_g, n3_g, x1_e, x2_n, x3_t, u10)
        st.dx = (s.dx1, s.dx2, s.dx3)
        return st

    # ---- harness file format (oracle/ref_harness.cpp) -------------------------
    @classmethod
    def from_dir(cls, path: str) -> "OracleState":
        meta = read_meta(path)
        st = cls(meta["N1"], meta["N2"], meta["N3"], dx=(meta["dx1"], meta["dx2"], meta["dx3"]),
                 dt=meta["dt"], CFL=meta["CFL"], VIS=meta["VIS"], sound=meta["sound"],
                 ro0=meta["ro0"], t0=meta["t0"])
        for name in FIELDS:
            a = np.fromfile(os.path.join(path, name + ".f64"), dtype="<f8")
            st.arrays[name] = a.reshape(st.shape(name)).astype(np.float64)
        return st

    def to_dir(self, path: str) -> None:
        os.makedirs(path, exist_ok=True)
        for name in FIELDS:
            self.arrays[name].astype("<f8").tofile(os.path.join(path, name + ".f64"))
        with open(os.path.join(path, "meta.txt"), "w") as fh:
            fh.write(f"N1 {self.N[0]}\nN2 {self.N[1]}\nN3 {self.N[2]}\n")
            for k, v in (("dx1", self.dx[0]), ("dx2", self.dx[1]), ("dx3", self.dx[2]),
                         ("dt", self.dt), ("CFL", self.CFL), ("VIS", self.VIS),
                         ("sound", self.sound), ("ro0", self.ro0), ("t0", self.t0),
                         ("TIME", 0.0)):
                fh.write(f"{k} {float(v).hex()}\n")
            fh.write("nStep 0\n")

    # ---- npz fixtures -----------------------------------------------------------
    def scalars(self) -> Dict[str, float]:
        return {"N1": self.N[0], "N2": self.N[1], "N3": self.N[2],
                "dx1": self.dx[0], "dx2": self.dx[1], "dx3": self.dx[2], "dt": self.dt,
                "CFL": self.CFL, "VIS": self.VIS, "sound": self.sound, "ro0": self.ro0,
                "t0": self.t0}


def read_meta(path: str) -> Dict[str, float]:
    meta: Dict[str, float] = {}
    with open(os.path.join(path, "meta.txt")) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) != 2:
                continue
            k, v = parts
            if k in ("N1", "N2", "N3", "nStep"):
                meta[k] = int(v)
            else:
                meta[k] = float.fromhex(v) if ("0x" in v or "p" in v) else float(v)
    return meta


def state_from_npz(z, prefix: str) -> OracleState:
    """Rebuild a state stored by tests/golden/make_fixtures.py under ``prefix``."""
    sc = {k[len(prefix) + 1:]: z[k] for k in z.files if k.startswith(prefix + ":")}
    st = OracleState(int(sc["N1"]), int(sc["N2"]), int(sc["N3"]),
                     dx=(float(sc["dx1"]), float(sc["dx2"]), float(sc["dx3"])),
                     dt=float(sc["dt"]), CFL=float(sc["CFL"]), VIS=float(sc["VIS"]),
                     sound=float(sc["sound"]), ro0=float(sc["ro0"]), t0=float(sc["t0"]))
    for name in FIELDS:
        if name in sc:
            st.arrays[name] = np.ascontiguousarray(sc[name], dtype=np.float64)
    return st


# ---- BoundaryConditions and halo planes -------------------------------------
def boundary_conditions(arrays, axis: int, periodic: bool, t0: float) -> None:
    """BoundaryConditions::alongX{axis+1}() (source/BoundaryConditions.cpp:22-85)
    in place on five cell arrays [u1, u2, u3, ro, t], numpy shape
    (N3+2, N2+2, N1+2).  n = N+1 is the reference's node count; the loops run
    over the interior 1..N of the two other axes."""
    u1, u2, u3, ro, t = arrays
    N3, N2, N1 = (s - 2 for s in u1.shape)
    if axis == 0:                                   # alongX1 (22-42): period flag unused
        n = N1 + 1
        for a in arrays:
            a[1:N3 + 1, 1:N2 + 1, n] = a[1:N3 + 1, 1:N2 + 1, n - 1]
            a[1:N3 + 1, 1:N2 + 1, 0] = a[1:N3 + 1, 1:N2 + 1, 1]
        return

    def sl(q):                                      # plane q of the axis, interior of the others
        return (slice(1, N3 + 1), q, slice(1, N1 + 1)) if axis == 1 else (q, slice(1, N2 + 1), slice(1, N1 + 1))

    n = (N2 if axis == 1 else N3) + 1
    if periodic:                                    # 44-63 (x2), 65-85 (x3)
        for a in (u1, u2, u3, ro):
            a[sl(n)] = a[sl(1)]
        t[sl(n)] = ro[sl(1)] if axis == 2 else t[sl(1)]   # x3: t from ro (73)
        for a in arrays:
            a[sl(0)] = a[sl(n - 1)]
    else:
        for a in (u1, u2, u3):
            a[sl(n)] = 0.0
            a[sl(0)] = 0.0
        ro[sl(n)] = ro[sl(n - 1)]
        ro[sl(0)] = ro[sl(1)]
        t[sl(n)] = t0
        t[sl(0)] = t0


def pack_plane(arr: np.ndarray, N, axis: int, q: int) -> np.ndarray:
    """The reference's halo packing loop (e.g. source/main.cpp:619-630): plane
    q normal to `axis` over the interior 1..N of the two other axes, as an
    Arr2d whose elem(a, b) = buf[b*na + a] (source/Arr2d.cpp:14-16), a being
    the lower of the two other axes; returned with numpy shape (nb, na)."""
    N1, N2, N3 = N
    if axis == 0:
        return arr[1:N3 + 1, 1:N2 + 1, q].copy()
    if axis == 1:
        return arr[1:N3 + 1, q, 1:N1 + 1].copy()
    return arr[q, 1:N2 + 1, 1:N1 + 1].copy()
