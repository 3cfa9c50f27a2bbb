"""Python front end of the MI355X CABARET/ARK step (``lib/libark_amd.so``).

A thin ctypes layer over the C-ABI in ``include/ark/ark.h``; it mirrors the
reference's phase-by-phase time-step API (source/main.cpp:12-17, 80-104):
``time_step_size``, ``phase1``, ``stress_tensor``, ``use_forces``,
``phase2``, plus the fused ``step``.  Fields are numpy arrays in the
reference's Arr3d layout and extents (source/Arr3d.cpp:19-21,
source/main.cpp:198-277): shape (N3+2, N2+2, N1+2) for cell fields and
(N3+3, N2+3, N1+3) for face and stress fields, x fastest.

There is no CPU fallback: every compute call runs the HIP kernels, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Iterable, Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libark_amd.so")

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
FIELD_ID = {n: q for q, n in enumerate(FIELDS)}
FACE_FIRST = FIELD_ID["ro1"]

ARK_OK, ARK_ERR_INVALID, ARK_ERR_HIP, ARK_ERR_UNSUPPORTED, ARK_ERR_NONFINITE, ARK_ERR_STATE = range(6)
PREC_FP64, PREC_MIXED = 0, 1
SCHEME_REFERENCE, SCHEME_CORRECTED = 0, 1

EXPORTED_SYMBOLS = [
    "ark_create", "ark_destroy", "ark_last_error", "ark_upload", "ark_download", "ark_set_dt",
    "ark_get_dt", "ark_next_step", "ark_half_time", "ark_set_closure_cutoff", "ark_set_tile_shape", "ark_get_tile_shape", "ark_time_step_size", "ark_phase1", "ark_stress_tensor", "ark_use_forces",
    "ark_phase2", "ark_step", "ark_advance", "ark_step_phases", "ark_diagnostics", "ark_init_tgv",
    "ark_synchronize", "ark_get_stream", "ark_profile_enable", "ark_profile_read",
    "ark_profile_reset", "ark_get_plane", "ark_set_plane", "ark_boundary_conditions", "ark_device_bytes",
    "ark_batch_create", "ark_batch_destroy", "ark_batch_last_error", "ark_batch_step", "ark_batch_synchronize",
    "ark_batch_autotune", "ark_batch_set_config", "ark_batch_get_config",
    "ark_init_turbulence", "ark_spectrum", "ark_dissipation", "ark_set_forcing", "ark_filter_box",
    "ark_mlp_param_count", "ark_mlp_create", "ark_mlp_destroy", "ark_mlp_last_error", "ark_mlp_set_params",
    "ark_mlp_get_params", "ark_mlp_set_normalization", "ark_mlp_set_data", "ark_mlp_gradient", "ark_mlp_train",
    "ark_mlp_predict", "ark_set_closure", "ark_init_velocity",
    "ark_write_paraview", "ark_average_values", "ark_write_average_values", "ark_init_channel",
    "ark_get_clock", "ark_restart", "ark_diagnostics_noalloc", "ark_init_state", "ark_filter_density",
]


class ArkError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ark status {status}: {message}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [
        ("N1", ctypes.c_int), ("N2", ctypes.c_int), ("N3", ctypes.c_int),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double), ("dx3", ctypes.c_double),
        ("CFL", ctypes.c_double), ("VIS", ctypes.c_double), ("sound", ctypes.c_double),
        ("ro0", ctypes.c_double), ("t0", ctypes.c_double),
        ("periodic", ctypes.c_int * 3),
        ("precision", ctypes.c_int), ("device", ctypes.c_int), ("scheme", ctypes.c_int),
    ]


class Diag(ctypes.Structure):
    _fields_ = [
        ("ke_reference", ctypes.c_double), ("ke_physical", ctypes.c_double),
        ("mass", ctypes.c_double), ("max_abs_u", ctypes.c_double),
        ("nonfinite", ctypes.c_int64), ("dt", ctypes.c_double), ("time", ctypes.c_double),
        ("nstep", ctypes.c_int64),
    ]


class Clock(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("nstep", ctypes.c_int64), ("dt", ctypes.c_double),
        ("dt_next", ctypes.c_double), ("dt_next_known", ctypes.c_int),
    ]


# The persistent state of a fluid, periodic box: what a checkpoint holds
# besides the clock (the face densities are derived from p on restart).
PERSISTENT = ["ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
              "u11", "u21", "u31", "p1", "t1", "u12", "u22", "u32", "p2", "t2",
              "u13", "u23", "u33", "p3", "t3"]

_lib = None


def build(force: bool = True) -> None:
    """Compile the library for gfx950 (hipcc cross-compiles without a GPU)."""
    args = ["make", "-s", "-C", PKG_DIR, "-j8"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


def lib() -> ctypes.CDLL:
    """Load libark_amd.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                           "(the HIP extension is required; there is no CPU path)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.ark_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.ark_destroy.argtypes = [vp]
    L.ark_destroy.restype = None
    L.ark_last_error.argtypes = [vp]
    L.ark_last_error.restype = ctypes.c_char_p
    L.ark_upload.argtypes = [vp, ip, dp, ip, ip, ip]
    L.ark_download.argtypes = [vp, ip, dp, ip, ip, ip]
    L.ark_set_dt.argtypes = [vp, ctypes.c_double]
    L.ark_get_dt.argtypes = [vp, dp]
    L.ark_time_step_size.argtypes = [vp, dp]
    for fn in ("ark_phase1", "ark_stress_tensor", "ark_use_forces", "ark_phase2", "ark_synchronize",
               "ark_profile_reset", "ark_next_step", "ark_half_time"):
        getattr(L, fn).argtypes = [vp]
    L.ark_step.argtypes = [vp, ip]
    L.ark_step_phases.argtypes = [vp, ip]
    L.ark_advance.argtypes = [vp, ip]
    L.ark_diagnostics.argtypes = [vp, ctypes.POINTER(Diag)]
    L.ark_init_tgv.argtypes = [vp, ctypes.c_double]
    L.ark_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.ark_profile_enable.argtypes = [vp, ip]
    L.ark_profile_read.argtypes = [vp, ip, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), dp,
                                   ctypes.POINTER(ctypes.c_int)]
    L.ark_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.ark_get_plane.argtypes = [vp, ip, ip, ip, dp, ip, ip]
    L.ark_set_plane.argtypes = [vp, ip, ip, ip, dp, ip, ip]
    L.ark_boundary_conditions.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int), ip, ctypes.c_double]
    L.ark_batch_create.argtypes = [ctypes.POINTER(vp), ip, ctypes.POINTER(vp)]
    L.ark_batch_destroy.argtypes = [vp]
    L.ark_batch_destroy.restype = None
    L.ark_batch_last_error.argtypes = [vp]
    L.ark_batch_last_error.restype = ctypes.c_char_p
    L.ark_batch_step.argtypes = [vp, ip]
    L.ark_batch_synchronize.argtypes = [vp]
    L.ark_batch_autotune.argtypes = [vp, ip, dp, ip]
    L.ark_batch_set_config.argtypes = [vp, ip, ip]
    L.ark_batch_get_config.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 4
    u64 = ctypes.c_ulonglong
    L.ark_init_turbulence.argtypes = [vp, u64, ctypes.c_double, ctypes.c_double]
    L.ark_spectrum.argtypes = [vp, ip, dp]
    L.ark_dissipation.argtypes = [vp, dp]
    L.ark_set_forcing.argtypes = [vp, u64, ctypes.c_double, ctypes.c_double]
    L.ark_filter_box.argtypes = [vp, ip, dp, dp, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    ll = ctypes.c_longlong
    L.ark_mlp_param_count.argtypes = []
    L.ark_mlp_create.argtypes = [ip, u64, ctypes.POINTER(vp)]
    L.ark_mlp_destroy.argtypes = [vp]
    L.ark_mlp_destroy.restype = None
    L.ark_mlp_last_error.argtypes = [vp]
    L.ark_mlp_last_error.restype = ctypes.c_char_p
    L.ark_mlp_set_params.argtypes = [vp, fp]
    L.ark_mlp_get_params.argtypes = [vp, fp]
    L.ark_mlp_set_normalization.argtypes = [vp, fp, fp, fp, fp]
    L.ark_mlp_set_data.argtypes = [vp, fp, fp, ll]
    L.ark_mlp_gradient.argtypes = [vp, fp, fp, ll, dp, dp]
    L.ark_mlp_train.argtypes = [vp, ip, ll, ctypes.c_float, dp]
    L.ark_mlp_predict.argtypes = [vp, fp, fp, ll]
    L.ark_set_closure.argtypes = [vp, vp, ctypes.c_double]
    L.ark_set_closure_cutoff.argtypes = [vp, ip]
    L.ark_set_tile_shape.argtypes = [vp, ip]
    L.ark_get_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.ark_write_paraview.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.ark_average_values.argtypes = [vp]
    L.ark_init_channel.argtypes = [vp, ctypes.c_double]
    L.ark_write_average_values.argtypes = [vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_int]
    L.ark_init_velocity.argtypes = [vp, dp]
    L.ark_init_state.argtypes = [vp, dp, dp]
    L.ark_filter_density.argtypes = [vp, ip, dp]
    L.ark_get_clock.argtypes = [vp, ctypes.POINTER(Clock)]
    L.ark_restart.argtypes = [vp, ctypes.POINTER(Clock)]
    L.ark_diagnostics_noalloc.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.c_void_p, ctypes.c_void_p]
    for fn in EXPORTED_SYMBOLS:
        if fn not in ("ark_destroy", "ark_last_error", "ark_batch_destroy", "ark_batch_last_error",
                      "ark_mlp_destroy", "ark_mlp_last_error"):
            getattr(L, fn).restype = ctypes.c_int
    _lib = L
    return L


def field_shape(name: str, N) -> tuple:
    N1, N2, N3 = N
    e = 3 if FIELD_ID[name] >= FACE_FIRST else 2
    return (N3 + e, N2 + e, N1 + e)


class Solver:
    """One rank's subdomain resident on one GPU (an ``ark_ctx``)."""

    def __init__(self, N1: int, N2: int, N3: int, dx=None, CFL: float = 0.3,
                 VIS: float = 1.0 / 1600.0, sound: float = 10.0, ro0: float = 1.0, t0: float = 1.0,
                 precision: int = PREC_FP64, device: int = 0, scheme: int = SCHEME_REFERENCE,
                 periodic=(1, 1, 1)):
        """periodic[d] = 0 makes axis d non-periodic as one rank of the
        reference with periods={0,0,0} sees it: its exchange delivers zeros
        and the reference's physical-boundary branches run (zero-gradient ends
        on X1 and X3, inflow and outflow on X1; walls come from `condition`)."""
        self.N = (int(N1), int(N2), int(N3))
        if dx is None:
            two_pi = 2.0 * np.pi
            dx = ((two_pi - 0.0) / N1, (two_pi - 0.0) / N2, (two_pi - 0.0) / N3)
        self.dx = tuple(float(v) for v in dx)
        p = Params()
        p.N1, p.N2, p.N3 = self.N
        p.dx1, p.dx2, p.dx3 = self.dx
        p.CFL, p.VIS, p.sound, p.ro0, p.t0 = CFL, VIS, sound, ro0, t0
        for d in range(3):
            p.periodic[d] = 1 if periodic[d] else 0
        p.precision, p.device, p.scheme = precision, device, scheme
        self.params = p
        self._L = lib()
        h = ctypes.c_void_p()
        st = self._L.ark_create(ctypes.byref(p), ctypes.byref(h))
        self._h = h
        if st != ARK_OK:
            msg = self._L.ark_last_error(h).decode() if h.value else "ark_create failed"
            self.close()
            raise ArkError(st, msg)

    # ---- plumbing -----------------------------------------------------------
    def _check(self, st: int) -> None:
        if st != ARK_OK:
            raise ArkError(st, self._L.ark_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ark_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ------------------------------------------------------------
    def shape(self, name: str) -> tuple:
        return field_shape(name, self.N)

    def upload(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        shp = self.shape(name)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape} != {shp}")
        self._check(self._L.ark_upload(self._h, FIELD_ID[name],
                                       a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       shp[2], shp[1], shp[0]))

    def download(self, name: str) -> np.ndarray:
        shp = self.shape(name)
        out = np.empty(shp, dtype=np.float64)
        self._check(self._L.ark_download(self._h, FIELD_ID[name],
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         shp[2], shp[1], shp[0]))
        return out

    def upload_all(self, arrays: Dict[str, np.ndarray], names: Optional[Iterable[str]] = None) -> None:
        for n in (names or FIELDS):
            if n in arrays:
                self.upload(n, arrays[n])

    def download_all(self, names: Optional[Iterable[str]] = None) -> Dict[str, np.ndarray]:
        return {n: self.download(n) for n in (names or FIELDS)}

    @property
    def dt(self) -> float:
        v = ctypes.c_double()
        self._check(self._L.ark_get_dt(self._h, ctypes.byref(v)))
        return v.value

    @dt.setter
    def dt(self, value: float) -> None:
        self._check(self._L.ark_set_dt(self._h, float(value)))

    # ---- the reference's phases (source/main.cpp:374-2495) --------------------
    def time_step_size(self) -> float:
        v = ctypes.c_double()
        self._check(self._L.ark_time_step_size(self._h, ctypes.byref(v)))
        return v.value

    def phase1(self) -> None:
        self._check(self._L.ark_phase1(self._h))

    def stress_tensor(self) -> None:
        self._check(self._L.ark_stress_tensor(self._h))

    def use_forces(self) -> None:
        self._check(self._L.ark_use_forces(self._h))

    def phase2(self) -> None:
        self._check(self._L.ark_phase2(self._h))

    def phase(self, name: str) -> None:
        {"TimeStepSize": self.time_step_size, "Phase1": self.phase1,
         "StressTensor": self.stress_tensor, "UseForces": self.use_forces,
         "Phase2": self.phase2, "Step": lambda: self.step(1)}[name]()

    def next_step(self) -> None:
        """nStep++ of the reference's loop (source/main.cpp:83)."""
        self._check(self._L.ark_next_step(self._h))

    def half_time(self) -> None:
        """TIME += 0.5 * dt of the reference's loop (source/main.cpp:84, 91)."""
        self._check(self._L.ark_half_time(self._h))

    def step(self, nsteps: int = 1) -> None:
        """nsteps of the driver loop (source/main.cpp:82-91), fused kernels."""
        self._check(self._L.ark_step(self._h, int(nsteps)))

    def advance(self, nsteps: int = 1) -> None:
        """nsteps fused steps maintaining only the persistent state (ark_advance)."""
        self._check(self._L.ark_advance(self._h, int(nsteps)))

    def step_phases(self, nsteps: int = 1) -> None:
        """nsteps through one launch per reference phase."""
        self._check(self._L.ark_step_phases(self._h, int(nsteps)))

    # ---- the reference's output path (source/main.cpp:2577-2891) ----------------
    def write_paraview(self, dir_path: str, need_swap: bool = True) -> str:
        """WriteDataParaView: <dir_path>out_<nStep>.vtk; returns the path."""
        if not dir_path.endswith(os.sep):
            dir_path += os.sep
        self._check(self._L.ark_write_paraview(self._h, dir_path.encode(), 1 if need_swap else 0))
        return os.path.join(dir_path, f"out_{self.diagnostics(allow_nonfinite=True)['nstep']}.vtk")

    def average_values(self) -> None:
        """AverageValues: averageU += u * dt (device)."""
        self._check(self._L.ark_average_values(self._h))

    def write_average_values(self, dir_path: str, start_time: float, need_swap: bool = True) -> str:
        """WriteAverageValues: <dir_path>average.vtk; returns the path."""
        if not dir_path.endswith(os.sep):
            dir_path += os.sep
        self._check(self._L.ark_write_average_values(self._h, dir_path.encode(), float(start_time),
                                                     1 if need_swap else 0))
        return os.path.join(dir_path, "average.vtk")

    def init_channel(self, u10: float = 1.0) -> None:
        """The reference's own problem (InitializeData, source/main.cpp:280-372)."""
        self._check(self._L.ark_init_channel(self._h, float(u10)))

    @classmethod
    def channel(cls, n1_g: int = 201, n2_g: int = 61, n3_g: int = 2, x1_e: float = 20.0, x2_n: float = 6.0,
                x3_t: float = 0.1, u10: float = 1.0, VIS: float = 1.0 / 1000.0, CFL: float = 0.2,
                sound: float = 10.0, device: int = 0) -> "Solver":
        """A context holding the reference's own problem; the defaults are its
        Input() values (source/main.cpp:118-195), periods={0,0,0}."""
        dx = ((x1_e - 0.) / (n1_g - 1), (x2_n - 0.) / (n2_g - 1), (x3_t - 0.) / (n3_g - 1))
        s = cls(n1_g - 1, n2_g - 1, n3_g - 1, dx=dx, CFL=CFL, VIS=VIS, sound=sound, device=device,
                periodic=(0, 0, 0))
        s.init_channel(u10)
        return s

    def init_tgv(self, U0: float = 1.0) -> None:
        self._check(self._L.ark_init_tgv(self._h, float(U0)))

    def diagnostics(self, allow_nonfinite: bool = False) -> Dict[str, float]:
        d = Diag()
        st = self._L.ark_diagnostics(self._h, ctypes.byref(d))
        if st != ARK_OK and not (allow_nonfinite and st == ARK_ERR_NONFINITE):
            self._check(st)
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def synchronize(self) -> None:
        self._check(self._L.ark_synchronize(self._h))

    # ---- checkpoint and restart ---------------------------------------------
    def clock(self) -> Dict[str, float]:
        """TIME, nStep, dt and the next step's dt if already reduced."""
        k = Clock()
        self._check(self._L.ark_get_clock(self._h, ctypes.byref(k)))
        return {n: getattr(k, n) for n, _ in Clock._fields_}

    def restart(self, clock: Dict[str, float]) -> None:
        """Continue from uploaded arrays with the given clock (ark_restart)."""
        k = Clock()
        for n, _ in Clock._fields_:
            setattr(k, n, clock[n])
        self._check(self._L.ark_restart(self._h, ctypes.byref(k)))

    def checkpoint(self) -> Dict[str, object]:
        """The persistent state and the clock, enough to continue bit for bit
        (plus u1Con..u3Con when the next step's dt is still to be computed
        from them)."""
        clock = self.clock()
        names = PERSISTENT if clock["dt_next_known"] else PERSISTENT + ["u1Con", "u2Con", "u3Con"]
        return {"arrays": self.download_all(names), "clock": clock}

    def resume(self, ckpt: Dict[str, object]) -> None:
        self.upload_all(ckpt["arrays"], list(ckpt["arrays"]))
        self.restart(ckpt["clock"])

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(self._L.ark_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    # ---- homogeneous turbulence ------------------------------------------------
    def init_turbulence(self, seed: int, k0: float = 4.0, urms: float = 1.0) -> None:
        self._check(self._L.ark_init_turbulence(self._h, int(seed), float(k0), float(urms)))

    def spectrum(self, nbins: Optional[int] = None) -> np.ndarray:
        nb = nbins or (int(np.ceil(np.sqrt(3) * max(self.N) / 2)) + 2)
        out = np.zeros(nb)
        self._check(self._L.ark_spectrum(self._h, nb, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def dissipation(self) -> float:
        v = ctypes.c_double()
        self._check(self._L.ark_dissipation(self._h, ctypes.byref(v)))
        return v.value

    def set_closure(self, mlp, delta: float) -> None:
        """Attach a trained MLP (or None) as the subgrid closure."""
        self._closure = mlp
        self._check(self._L.ark_set_closure(self._h, mlp._h if mlp is not None else None, float(delta)))

    @property
    def tile_shape(self) -> int:
        """Mixed-precision tile shape: -1 (chosen by timing at the next step),
        0 (30x14 owned) or 1 (32x13); results are the same for every shape."""
        v = ctypes.c_int()
        self._check(self._L.ark_get_tile_shape(self._h, ctypes.byref(v)))
        return v.value

    @tile_shape.setter
    def tile_shape(self, shape: int) -> None:
        self._check(self._L.ark_set_tile_shape(self._h, int(shape)))

    def set_closure_cutoff(self, kcut: int) -> None:
        """Apply only the shells round(|m|) >= kcut of the closure's
        acceleration (ark_set_closure_cutoff); 0 applies all of it."""
        self._check(self._L.ark_set_closure_cutoff(self._h, int(kcut)))

    def init_velocity(self, u) -> None:
        """Cells from u (3, N3, N2, N1); faces by the half-cell spectral shift."""
        N1, N2, N3 = self.N
        a = np.ascontiguousarray(u, dtype=np.float64)
        if a.shape != (3, N3, N2, N1):
            raise ValueError(f"velocity shape {a.shape} != {(3, N3, N2, N1)}")
        self._check(self._L.ark_init_velocity(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def init_state(self, u, rho) -> None:
        """init_velocity plus the density rho (N3, N2, N1) on the cells and,
        by the same shift, on the faces with their pressures."""
        N1, N2, N3 = self.N
        a = np.ascontiguousarray(u, dtype=np.float64)
        r = np.ascontiguousarray(rho, dtype=np.float64)
        if a.shape != (3, N3, N2, N1) or r.shape != (N3, N2, N1):
            raise ValueError(f"state shapes {a.shape}, {r.shape} != {(3, N3, N2, N1)}, {(N3, N2, N1)}")
        P = ctypes.POINTER(ctypes.c_double)
        self._check(self._L.ark_init_state(self._h, a.ctypes.data_as(P), r.ctypes.data_as(P)))

    def filter_density(self, width: int) -> np.ndarray:
        """Block means of the cell density over width^3 cells, (Nc3, Nc2, Nc1)."""
        N1, N2, N3 = self.N
        out = np.zeros((N3 // width, N2 // width, N1 // width))
        self._check(self._L.ark_filter_density(self._h, int(width), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_forcing(self, seed: int, eps: float, tau: float = 0.0) -> None:
        self._check(self._L.ark_set_forcing(self._h, int(seed), float(eps), float(tau)))

    def filter_box(self, width: int):
        """(ubar[3], tau[6], grad[9]) on the coarse grid, numpy (Nc3, Nc2, Nc1)."""
        N1, N2, N3 = self.N
        shp = (N3 // width, N2 // width, N1 // width)
        ub, tau, gr = np.zeros((3,) + shp), np.zeros((6,) + shp), np.zeros((9,) + shp)
        P = ctypes.POINTER(ctypes.c_double)
        self._check(self._L.ark_filter_box(self._h, int(width), ub.ctypes.data_as(P), tau.ctypes.data_as(P),
                                           gr.ctypes.data_as(P)))
        return ub, tau, gr

    # ---- halo planes and BoundaryConditions -----------------------------------
    def plane_shape(self, axis: int) -> tuple:
        """(nb, na) of a plane normal to `axis`: numpy shape of the Arr2d buffer."""
        N1, N2, N3 = self.N
        na = N2 if axis == 0 else N1
        nb = N2 if axis == 2 else N3
        return (nb, na)

    def get_plane(self, name: str, axis: int, q: int) -> np.ndarray:
        nb, na = self.plane_shape(axis)
        out = np.empty((nb, na), dtype=np.float64)
        self._check(self._L.ark_get_plane(self._h, FIELD_ID[name], axis, q,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), na, nb))
        return out

    def set_plane(self, name: str, axis: int, q: int, plane: np.ndarray) -> None:
        nb, na = self.plane_shape(axis)
        a = np.ascontiguousarray(plane, dtype=np.float64)
        if a.shape != (nb, na):
            raise ValueError(f"plane shape {a.shape} != {(nb, na)}")
        self._check(self._L.ark_set_plane(self._h, FIELD_ID[name], axis, q,
                                          a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), na, nb))

    def boundary_conditions(self, axis: int, fields, periodic: bool, t0: float) -> None:
        """BoundaryConditions::alongX{axis+1} on the fields (u1, u2, u3, ro, t)."""
        ids = (ctypes.c_int * 5)(*[FIELD_ID[f] for f in fields])
        self._check(self._L.ark_boundary_conditions(self._h, axis, ids, int(bool(periodic)), float(t0)))

    def device_bytes(self) -> int:
        v = ctypes.c_int64()
        self._check(self._L.ark_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    # ---- per-kernel timing (HIP events on the context's stream) ---------------
    def profile(self, on: bool = True) -> None:
        self._check(self._L.ark_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self) -> None:
        self._check(self._L.ark_profile_reset(self._h))

    def profile_read(self) -> Dict[str, tuple]:
        n = ctypes.c_int()
        cap = 64
        names = (ctypes.c_char * 64 * cap)()
        launches = (ctypes.c_int64 * cap)()
        total = (ctypes.c_double * cap)()
        self._check(self._L.ark_profile_read(self._h, cap, ctypes.cast(names, ctypes.c_void_p), launches,
                                             total, ctypes.byref(n)))
        out = {}
        for q in range(min(n.value, cap)):
            out[bytes(names[q]).split(b"\0", 1)[0].decode()] = (int(launches[q]), float(total[q]))
        return out


BATCH_SHAPES = ["30x6", "30x14", "62x6", "32x13"]
BATCH_LZ = [16, 32, 64]


class Batch:
    """Mixed-precision boxes advanced together by one launch per step
    (``ark_batch``).  The Solver objects stay usable and must outlive it."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        self._L = lib()
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        st = self._L.ark_batch_create(arr, len(self.solvers), ctypes.byref(h))
        if st != ARK_OK:
            raise ArkError(st, "ark_batch_create failed (boxes must be mixed precision, fluid, one device and "
                               "scheme, N >= 2)")
        self._h = h

    def _check(self, st: int) -> None:
        if st != ARK_OK:
            raise ArkError(st, self._L.ark_batch_last_error(self._h).decode())

    def step(self, nsteps: int = 1) -> None:
        self._check(self._L.ark_batch_step(self._h, int(nsteps)))

    def synchronize(self) -> None:
        self._check(self._L.ark_batch_synchronize(self._h))

    def autotune(self, reps: int = 3) -> Dict[tuple, float]:
        n = 3 * len(BATCH_SHAPES)
        ms = (ctypes.c_double * n)()
        self._check(self._L.ark_batch_autotune(self._h, int(reps), ms, n))
        return {(BATCH_SHAPES[i // 3], BATCH_LZ[i % 3]): ms[i] for i in range(n)}

    def set_config(self, shape: int, lz: int) -> None:
        self._check(self._L.ark_batch_set_config(self._h, int(shape), int(lz)))

    def config(self) -> Dict[str, int]:
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self._L.ark_batch_get_config(self._h, *[ctypes.byref(x) for x in v]))
        return {"shape": BATCH_SHAPES[v[0].value], "lz": v[1].value, "tiles": v[2].value, "threads": v[3].value}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ark_batch_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"shape {a.shape} != {shape}")
    return a


class MLP:
    """The closure network (``ark_mlp``): 10 -> 64 -> 64 -> 6, SiLU, BF16-MFMA."""

    NI, NO = 10, 6

    def __init__(self, seed: int = 0, device: int = 0):
        self._L = lib()
        self.nparam = self._L.ark_mlp_param_count()
        h = ctypes.c_void_p()
        st = self._L.ark_mlp_create(int(device), int(seed), ctypes.byref(h))
        if st != ARK_OK:
            raise ArkError(st, "ark_mlp_create failed")
        self._h = h

    def _check(self, st: int) -> None:
        if st != ARK_OK:
            raise ArkError(st, self._L.ark_mlp_last_error(self._h).decode())

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparam, dtype=np.float32)
        self._check(self._L.ark_mlp_get_params(self._h, self._p(out)))
        return out

    def set_params(self, p) -> None:
        p = _f32(p, (self.nparam,))
        self._check(self._L.ark_mlp_set_params(self._h, self._p(p)))

    def set_normalization(self, in_mean, in_std, out_mean, out_std) -> None:
        arrs = [_f32(in_mean, (10,)), _f32(in_std, (10,)), _f32(out_mean, (6,)), _f32(out_std, (6,))]
        self._check(self._L.ark_mlp_set_normalization(self._h, *[self._p(a) for a in arrs]))

    def set_data(self, X, T) -> None:
        X, T = _f32(X), _f32(T)
        self._check(self._L.ark_mlp_set_data(self._h, self._p(X), self._p(T), X.shape[0]))

    def gradient(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.zeros(self.nparam)
        loss = ctypes.c_double()
        self._check(self._L.ark_mlp_gradient(self._h, self._p(X), self._p(T), X.shape[0],
                                             g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(loss)))
        return g, loss.value

    def train(self, nsteps: int, batch: int = 65536, lr: float = 1e-3) -> float:
        loss = ctypes.c_double()
        self._check(self._L.ark_mlp_train(self._h, int(nsteps), int(batch), float(lr), ctypes.byref(loss)))
        return loss.value

    def predict(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], 6), dtype=np.float32)
        self._check(self._L.ark_mlp_predict(self._h, self._p(X), self._p(Y), X.shape[0]))
        return Y

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ark_mlp_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
