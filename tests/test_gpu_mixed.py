"""Mixed-precision mode (FP32 fluxes/invariants/stress, FP64 closing update
and global sums) against the FP64 path, which is itself bitwise equal to the
reference.  Tolerances are written per test.
"""
import numpy as np
import pytest

from conftest import bits_equal
import arkamd

pytestmark = pytest.mark.gpu

CELL = ["ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon"]
FACE = ["u11", "u21", "u31", "p1", "t1", "u12", "u22", "u32", "p2", "t2", "u13", "u23", "u33", "p3", "t3"]


def interior(name, a, N):
    N1, N2, N3 = N
    if name.endswith("Con"):
        return a[1:N3 + 1, 1:N2 + 1, 1:N1 + 1]
    d = int(name[-1]) - 1
    e = [N1 + (d == 0), N2 + (d == 1), N3 + (d == 2)]
    return a[1:e[2] + 1, 1:e[1] + 1, 1:e[0] + 1]


def run_pair(N, steps, scheme=arkamd.SCHEME_CORRECTED):
    out = {}
    for prec in (arkamd.PREC_FP64, arkamd.PREC_MIXED):
        with arkamd.Solver(*N, precision=prec, scheme=scheme) as s:
            s.init_tgv()
            s.step(steps)
            out[prec] = ({n: interior(n, s.download(n), N) for n in CELL + FACE}, s.diagnostics())
    return out


def rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_mixed_init_roundtrip():
    N = (16, 12, 10)
    with arkamd.Solver(*N) as d, arkamd.Solver(*N, precision=arkamd.PREC_MIXED) as m:
        d.init_tgv()
        m.init_tgv()
        for n in CELL:   # FP64 cell storage: identical
            assert np.array_equal(d.download(n), m.download(n)), n
        for n in FACE:   # FP32 face storage: FP64 values rounded once
            a, b = d.download(n), m.download(n)
            if n.startswith("t"):
                assert np.array_equal(a, b)          # t' = 0 is exact
            else:
                assert np.array_equal(a.astype(np.float32).astype(np.float64), b), n
        # face densities are derived from p
        ro = m.download("ro2")
        p = m.download("p2")
        assert np.allclose(ro[1:-1, 1:-1, 1:-1], 1.0 + p[1:-1, 1:-1, 1:-1] / 100.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("N,steps,tol", [((32, 32, 32), 100, 1e-5), ((24, 20, 16), 60, 1e-5),
                                         ((70, 40, 90), 40, 1e-5), ((4, 6, 4), 20, 1e-5)])
def test_mixed_close_to_fp64(N, steps, tol):
    """Relative L2 error per field after `steps` steps of the corrected scheme
    (the reference scheme goes unstable on this vortex, after which any
    rounding difference grows to O(1)); tolerance tol on velocities and
    density, looser on p and rho - 1, which are Ma^2 smaller than u."""
    r = run_pair(N, steps)
    f64, dmix = r[arkamd.PREC_FP64], r[arkamd.PREC_MIXED]
    errs = {n: rel_l2(dmix[0][n], f64[0][n]) for n in CELL + FACE}
    for n in ("ronCon", "u1nCon", "u2nCon", "u11", "u22"):
        assert errs[n] < tol, (n, errs[n])
    for n in ("p1", "p2", "p3"):
        assert errs[n] < 1e-3, (n, errs[n])
    # the density perturbation rho - 1 itself
    dr_err = rel_l2(dmix[0]["ronCon"] - 1.0, f64[0]["ronCon"] - 1.0)
    assert dr_err < 1e-3, dr_err
    assert abs(dmix[1]["dt"] - f64[1]["dt"]) / f64[1]["dt"] < 1e-6
    assert dmix[1]["nstep"] == f64[1]["nstep"] == steps


def test_mixed_odd_box_within_fp64_sensitivity():
    """On a box with odd sides the Taylor-Green vortex loses the grid
    symmetry that keeps it regular, and a few cells per axis make the step's
    limiter and upwind branches flip on rounding-size differences: FP64 runs
    whose initial velocities differ by 1e-7 relative part by 1e-3 to 1e-1
    within 20 steps.  The mixed run must stay inside that envelope."""
    N, steps = (6, 5, 7), 20
    r = run_pair(N, steps)
    f64, mix = r[arkamd.PREC_FP64][0], r[arkamd.PREC_MIXED][0]
    rng = np.random.default_rng(0)
    with arkamd.Solver(*N, scheme=arkamd.SCHEME_CORRECTED) as s:
        s.init_tgv()
        for n in ("u1nCon", "u2nCon"):
            a = s.download(n)
            s.upload(n, a * (1.0 + 1e-7 * rng.standard_normal(a.shape)))
        s.step(steps)
        pert = {n: interior(n, s.download(n), N) for n in ("u1nCon", "u2nCon")}
    for n in ("u1nCon", "u2nCon"):
        e_mix, e_pert = rel_l2(mix[n], f64[n]), rel_l2(pert[n], f64[n])
        assert e_mix < 3 * e_pert, (n, e_mix, e_pert)


def test_mixed_mass_conservation():
    with arkamd.Solver(32, 32, 32, precision=arkamd.PREC_MIXED) as s:
        s.init_tgv()
        m0 = s.diagnostics()["mass"]
        s.step(40)
        m1 = s.diagnostics()["mass"]
    assert abs(m1 - m0) / m0 < 1e-9


def test_mixed_uniform_state_stays_uniform():
    N = (12, 12, 12)
    with arkamd.Solver(*N, precision=arkamd.PREC_MIXED) as s:
        for n in arkamd.FIELDS:
            shp = s.shape(n)
            if n in ("ronCon", "roCon") or n.startswith("ro"):
                s.upload(n, np.ones(shp))
            elif n.startswith("t"):
                s.upload(n, np.ones(shp))
            elif n in ("u1nCon", "u1Con") or n in ("u11", "u12", "u13"):
                s.upload(n, np.full(shp, 0.5))
            else:
                s.upload(n, np.zeros(shp))
        s.step(5)
        u = interior("u1nCon", s.download("u1nCon"), N)
        ro = interior("ronCon", s.download("ronCon"), N)
        assert np.abs(u - 0.5).max() < 1e-6 and np.abs(ro - 1.0).max() < 1e-6


def test_mixed_rejects_phase_calls():
    with arkamd.Solver(8, 8, 8, precision=arkamd.PREC_MIXED) as s:
        s.init_tgv()
        with pytest.raises(arkamd.ArkError) as e:
            s.phase1()
        assert e.value.status == arkamd.ARK_ERR_UNSUPPORTED
        s.step(2)   # the fused mixed path works on the same context
        assert s.diagnostics()["nstep"] == 2


def test_mixed_reference_scheme_before_instability():
    """The reference scheme too, over a stretch where it is still stable."""
    r = run_pair((32, 32, 32), 30, scheme=arkamd.SCHEME_REFERENCE)
    f64, dmix = r[arkamd.PREC_FP64], r[arkamd.PREC_MIXED]
    for n in ("ronCon", "u1nCon", "u2nCon", "u11", "u22"):
        assert rel_l2(dmix[0][n], f64[0][n]) < 1e-5, n


@pytest.mark.parametrize("scheme", [arkamd.SCHEME_CORRECTED, arkamd.SCHEME_REFERENCE])
@pytest.mark.parametrize("kind", ["tgv", "forced"])
def test_mixed_tile_shapes_bitwise_equal(kind, scheme):
    """The single-box step's tile shapes (30x14 and 32x13 owned columns)
    compute every cell the same way: forced shapes give the same state bit
    for bit, and a context left to time them takes one of the two."""
    N = (70, 46, 34)   # even sides (the turbulence initialiser's transforms); ragged against both tilings
    outs = []
    for shape in (0, 1, -1):
        with arkamd.Solver(*N, VIS=1 / 800, precision=arkamd.PREC_MIXED, scheme=scheme) as s:
            if kind == "tgv":
                s.init_tgv(1.0)
            else:
                s.init_turbulence(7, 3.0, 1.0)
                s.set_forcing(8, 0.1, 0.0)
            s.tile_shape = shape
            s.step(6)
            assert s.tile_shape in (0, 1) and (shape < 0 or s.tile_shape == shape)
            outs.append({n: s.download(n) for n in ("ronCon", "u1nCon", "u3nCon", "tnCon", "p1", "u22", "t3")})
    for n in outs[0]:
        assert bits_equal(outs[0][n], outs[1][n]), n
        assert bits_equal(outs[0][n], outs[2][n]), n


# ---------------------------------------------------------------------------
# The 256^3 Taylor-Green workload itself, every persistent field.
def _mixed_box(kind, N=(24, 20, 16)):
    s = arkamd.Solver(*N, precision=arkamd.PREC_MIXED, scheme=arkamd.SCHEME_CORRECTED)
    if kind == "tgv":
        s.init_tgv(1.0)
    else:
        s.init_turbulence(5, 4.0, 1.0)
        s.set_forcing(6, 0.1, 0.01)
    return s


def _periodic_images(a, N):
    """A cell array (extents N + 2) rebuilt from its interior: ghost 0 is
    cell N and ghost N + 1 is cell 1 along every axis."""
    N1, N2, N3 = N
    return np.take(np.take(np.take(a, [N3, *range(1, N3 + 1), 1], 0), [N2, *range(1, N2 + 1), 1], 1),
                   [N1, *range(1, N1 + 1), 1], 2)


@pytest.mark.parametrize("kind", ["tgv", "forced"])
def test_mixed_cell_layout_changes_keep_steps_bitwise(kind):
    """Mixed steps hold the level-n cells as 32-byte records; every per-field
    reader (download, diagnostics, spectrum) converts them back first, and the
    next step converts again.  Five steps taken with such reads between them
    equal five steps in one call bit for bit, every persistent array whole,
    and the cell ghosts are the periodic images of the interior (the step
    itself reads no ghost)."""
    N = (24, 20, 16)
    with _mixed_box(kind, N) as a:
        a.step(5)
        ref = a.download_all(arkamd.PERSISTENT)
    with _mixed_box(kind, N) as b:
        b.step(2)
        b.download("u2nCon")
        b.diagnostics()
        b.step(1)
        b.spectrum(8)
        b.step(1)
        b.download("tnCon")
        b.step(1)
        got = b.download_all(arkamd.PERSISTENT)
    for n in arkamd.PERSISTENT:
        assert bits_equal(got[n], ref[n]), n
    for n in CELL:
        assert bits_equal(ref[n], _periodic_images(ref[n], N)), n


def test_batch_after_a_box_is_read_between_batch_steps():
    """A box read on its own between batch steps (its cells converted to the
    per-field layout and back, one buffer swap each way) still steps in the
    batch bit for bit like a standalone run: the batch rebuilds its argument
    tables when any buffer group of a box has swapped outside its steps."""
    kinds = ["tgv", "forced", "tgv"]
    alone = []
    for q, k in enumerate(kinds):
        with _mixed_box(k) as s:
            s.step(6 if q == 0 else 5)      # box 0 also takes one step on its own
            alone.append(s.download_all(arkamd.PERSISTENT))
    boxes = [_mixed_box(k) for k in kinds]
    try:
        with arkamd.Batch(boxes) as b:
            b.step(2)
            b.synchronize()
            boxes[1].download("u1nCon")
            b.step(2)
            b.synchronize()
            boxes[0].step(1)
            b.step(1)
            b.synchronize()
            boxes[0].download("ronCon")
        for q, s in enumerate(boxes):
            got = s.download_all(arkamd.PERSISTENT)
            for n in arkamd.PERSISTENT:
                assert bits_equal(got[n], alone[q][n]), (q, n)
    finally:
        for s in boxes:
            s.close()


VEL = {"u1nCon", "u2nCon", "u3nCon", "u11", "u21", "u31", "u12", "u22", "u32", "u13", "u23", "u33"}
PRES = {"p1", "p2", "p3"}
TEMP = {"tnCon", "t1", "t2", "t3"}


def field_errors(mix, f64):
    """Per-field relative L2 errors with a scale that stays meaningful where a
    field is near zero:
      velocities (cells and faces, every component): |u_mix - u_64| over the
        norm of the whole FP64 velocity vector at the same positions (u3 of
        the vortex starts at zero and stays small, so its own norm would
        measure the rounding of a near-zero field, not the error of the flow);
      density: |rho_mix - rho_64| / |rho_64|, and its perturbation rho - rho0
        against the FP64 perturbation;
      pressures: |p_mix - p_64| / |p_64|;
      t: the vortex carries t = t0, so the largest absolute difference."""
    cell_u = np.sqrt(sum(np.sum(f64[n] ** 2) for n in ("u1nCon", "u2nCon", "u3nCon")))
    face_u = {d: np.sqrt(sum(np.sum(f64[f"u{m}{d}"] ** 2) for m in (1, 2, 3))) for d in (1, 2, 3)}
    err = {}
    for n in VEL:
        scale = cell_u if n.endswith("Con") else face_u[int(n[-1])]
        err[n] = float(np.sqrt(np.sum((mix[n] - f64[n]) ** 2)) / scale)
    err["ronCon"] = rel_l2(mix["ronCon"], f64["ronCon"])
    err["ronCon-ro0"] = rel_l2(mix["ronCon"] - 1.0, f64["ronCon"] - 1.0)
    for n in PRES:
        err[n] = rel_l2(mix[n], f64[n])
    for n in TEMP:
        err[n] = float(np.abs(mix[n] - f64[n]).max())
    return err


def fp32_envelope(N, steps, scheme):
    """The FP64 step started from the initial state rounded once to FP32:
    how far the exact arithmetic itself moves each field when its input is
    perturbed by one FP32 rounding, which mixed precision's FP32 face storage
    does to every face every step."""
    with arkamd.Solver(*N, scheme=scheme) as s:
        s.init_tgv()
        for n in CELL + FACE:
            s.upload(n, s.download(n).astype(np.float32).astype(np.float64))
        s.step(steps)
        return {n: interior(n, s.download(n), N) for n in CELL + FACE}


def check_against_fp64(N, steps, scheme):
    """Mixed against FP64 after `steps` steps of the 256^3 workload, all 20
    persistent fields.  The solution -- the cell velocities and density --
    must meet the bound rel-L2 <= 1e-5.  Every field, faces included, must
    stay within 1.25x of the FP32 envelope (fp32_envelope): the face
    pressures (about 2e-4 after 500 steps) and the tangential face
    velocities (about 1e-4, chosen by the sign of the normal velocity of two
    cells, which flips where that velocity is near zero) are that sensitive
    to one rounding of their input in FP64 too, so no FP32 storage of them
    can be closer.  t is held to 1e-12 absolute."""
    r = run_pair(N, steps, scheme=scheme)
    f64, mix = r[arkamd.PREC_FP64], r[arkamd.PREC_MIXED]
    err = field_errors(mix[0], f64[0])
    env = field_errors(fp32_envelope(N, steps, scheme), f64[0])
    print(f"{N} x {steps} scheme {scheme}",
          {k: f"{err[k]:.2e}/{env[k]:.2e}" for k in sorted(err)})
    for n in ("u1nCon", "u2nCon", "u3nCon", "ronCon"):
        assert err[n] <= 1e-5, (n, err[n])
    for n in err:
        if n in TEMP:
            assert err[n] <= 1e-12, (n, err[n])
        else:
            assert err[n] <= 1.25 * env[n], (n, err[n], env[n])
    assert abs(mix[1]["dt"] - f64[1]["dt"]) / f64[1]["dt"] < 1e-5
    assert mix[1]["nstep"] == f64[1]["nstep"] == steps
    assert abs(mix[1]["ke_physical"] - f64[1]["ke_physical"]) / f64[1]["ke_physical"] < 1e-6
    return f64[1]


def test_mixed_256_every_field():
    """The corrected scheme over the workload's 500 steps (the reference
    scheme is unstable on this vortex; next test)."""
    check_against_fp64((256, 256, 256), 500, arkamd.SCHEME_CORRECTED)


def test_mixed_256_reference_scheme_stable_window():
    """The timed kernel (reference scheme, 32x13 tiles at 256^3) over 40
    steps.  Its unstable mode grows from rounding noise from the first step
    (FP64 runs whose inputs differ by one FP32 rounding part by 3e-5 in p
    after 10 steps and 1e-3 after 40) and is visible in max|u| after about
    100; within 40 steps the cells still meet 1e-5."""
    d = check_against_fp64((256, 256, 256), 40, arkamd.SCHEME_REFERENCE)
    assert d["max_abs_u"] < 1.001, d["max_abs_u"]


@pytest.mark.parametrize("scheme", [arkamd.SCHEME_REFERENCE, arkamd.SCHEME_CORRECTED])
def test_mixed_256_tile_shapes_bitwise(scheme):
    """At 256^3 (the 1024-chunk table, 32-bit byte offsets up to the top of
    the field group) the two single-box tilings give the same bits."""
    N = (256, 256, 256)
    outs = []
    for shape in (0, 1):
        with arkamd.Solver(*N, precision=arkamd.PREC_MIXED, scheme=scheme) as s:
            s.init_tgv()
            s.tile_shape = shape
            s.step(4)
            outs.append({n: s.download(n) for n in ("ronCon", "u1nCon", "u3nCon", "p1", "u22", "u33", "t3")})
    for n in outs[0]:
        assert bits_equal(outs[0][n], outs[1][n]), n


def test_mixed_vtk_vorticity_from_face_records(tmp_path):
    """ark_write_paraview on a mixed-precision context: its vorticity R is the
    reference's face-difference expression (source/main.cpp:2712-2733) on the
    FP32 face values, which download exactly as FP64; recomputed here in FP64
    from the downloaded faces, bit for bit."""
    N1, N2, N3 = 12, 10, 8
    with arkamd.Solver(N1, N2, N3, precision=arkamd.PREC_MIXED, scheme=arkamd.SCHEME_CORRECTED) as s:
        s.init_tgv()
        s.step(3)
        path = s.write_paraview(str(tmp_path))
        f = {n: s.download(n) for n in ("u12", "u13", "u21", "u23", "u31", "u32")}
        dx1, dx2, dx3 = s.dx
    data = open(path, "rb").read()
    at = data.index(b"VECTORS R double\n") + len(b"VECTORS R double\n")
    R = np.frombuffer(data[at:at + 3 * N1 * N2 * N3 * 8], dtype=">f8").reshape(N3, N2, N1, 3)
    I, J, K = slice(1, N1 + 1), slice(1, N2 + 1), slice(1, N3 + 1)
    Jp, Kp, Ip = slice(2, N2 + 2), slice(2, N3 + 2), slice(2, N1 + 2)
    d2u3 = (f["u32"][K, Jp, I] - f["u32"][K, J, I]) / dx2
    d3u2 = (f["u23"][Kp, J, I] - f["u23"][K, J, I]) / dx3
    d3u1 = (f["u13"][Kp, J, I] - f["u13"][K, J, I]) / dx3
    d1u3 = (f["u31"][K, J, Ip] - f["u31"][K, J, I]) / dx1
    d1u2 = (f["u21"][K, J, Ip] - f["u21"][K, J, I]) / dx1
    d2u1 = (f["u12"][K, Jp, I] - f["u12"][K, J, I]) / dx2
    want = np.stack([d2u3 - d3u2, d3u1 - d1u3, d1u2 - d2u1], axis=-1)
    assert bits_equal(R.astype(np.float64), want)
