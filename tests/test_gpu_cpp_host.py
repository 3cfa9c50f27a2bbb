"""GPU: the C++ host interface (include/ark/ark_host.hpp) through its two
programs.  ark_driver runs the reference's time loop (source/main.cpp:80-104)
and must reproduce the reference's 16^3 x 10 Taylor-Green digests bit for bit,
phase by phase and fused, and, with --channel, the reference's whole program
on its own problem with its output files; ark_host_check drives BoundaryConditions and Arr2d
exactly as a reference-side program would and must reproduce the reference
BoundaryConditions class and halo packing loops."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import ark_oracle as orc
from conftest import GOLDEN, bits_equal

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "..", "ark-cpp-mpi_amd", "bin")
DRIVER = os.path.join(BIN, "ark_driver")
CHECK = os.path.join(BIN, "ark_host_check")


@pytest.mark.parametrize("mode", [[], ["--phases"]])
def test_driver_reproduces_reference_digest(tmp_path, mode):
    run = json.load(open(os.path.join(GOLDEN, "tgv_digests.json")))["runs"]["16^3x10"]
    out = tmp_path / "state"
    res = subprocess.run([DRIVER] + mode + ["16", "16", "16", "10", str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    rec = json.loads(res.stdout.strip().splitlines()[-1])
    assert rec["nonfinite"] == 0
    assert rec["dt"] == float.fromhex(run["final_dt"])
    bad = []
    for name in orc.FIELDS:
        data = np.fromfile(out / f"{name}.f64", dtype="<f8")
        if hashlib.sha256(data.tobytes()).hexdigest() != run["arrays"][name]["sha256"]:
            bad.append(name)
    assert not bad, f"{mode}: arrays differing from the reference: {bad}"


@pytest.mark.parametrize("mode", [[], ["--phases"]])
def test_driver_runs_the_reference_program(tmp_path, mode):
    """ark_driver --channel is the reference's main() on its own problem: the
    first output, 20 steps with WriteDataParaView every 10 from step 10 and
    the running averages from step 5 -- the files the reference writes, byte
    for byte, and its progress lines."""
    files = json.load(open(os.path.join(GOLDEN, "channel_digests.json")))["vtk"]["files"]
    out = tmp_path / "out"
    res = subprocess.run([DRIVER, "--channel"] + mode + ["20", str(out), "10", "10", "5"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.strip().splitlines()
    assert [l.split()[:2] for l in lines[:-1]] == [["step:", "10"], ["step:", "20"]]
    rec = json.loads(lines[-1])
    assert rec["nstep"] == 20 and rec["nonfinite"] == 0
    assert (out / "out_0.vtk").exists()
    for name in ("out_10.vtk", "out_20.vtk", "average.vtk"):
        data = (out / name).read_bytes()
        assert len(data) == files[name]["bytes"], name
        assert hashlib.sha256(data).hexdigest() == files[name]["sha256"], name


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("periodic", [0, 1])
def test_cpp_boundary_conditions_vs_reference(tmp_path, axis, periodic):
    bc = np.load(os.path.join(GOLDEN, "bc.npz"))
    N1, N2, N3 = (int(v) for v in bc["N"])
    fin, fout = tmp_path / "in.f64", tmp_path / "out.f64"
    bc["input"].astype("<f8").tofile(fin)
    res = subprocess.run([CHECK, "bc", str(N1), str(N2), str(N3), str(axis), str(periodic), repr(float(bc["t0"])),
                          str(fin), str(fout)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    got = np.fromfile(fout, dtype="<f8").reshape(bc["input"].shape)
    assert bits_equal(got, bc[f"out_{axis}_{periodic}"])


@pytest.mark.parametrize("field,axis,q", [("u2nCon", 0, 1), ("u2nCon", 2, 7), ("p1", 1, 0), ("u13", 2, 9)])
def test_cpp_arr2d_pack_unpack(tmp_path, field, axis, q):
    N = (6, 5, 7)
    fid = orc.FIELDS.index(field)
    e = 3 if orc.is_face(field) else 2
    a = np.random.default_rng(fid).normal(size=(N[2] + e, N[1] + e, N[0] + e))
    fin, fout = tmp_path / "in.f64", tmp_path / "out.f64"
    a.astype("<f8").tofile(fin)
    res = subprocess.run([CHECK, "plane", *map(str, N), str(fid), str(axis), str(q), str(fin), str(fout)],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    raw = np.fromfile(fout, dtype="<f8")
    plane = orc.pack_plane(a, N, axis, q)
    assert bits_equal(raw[:plane.size].reshape(plane.shape), plane)
    b = a.copy()
    if axis == 0:
        b[1:N[2] + 1, 1:N[1] + 1, q] = -plane
    elif axis == 1:
        b[1:N[2] + 1, q, 1:N[0] + 1] = -plane
    else:
        b[q, 1:N[1] + 1, 1:N[0] + 1] = -plane
    assert bits_equal(raw[plane.size:].reshape(a.shape), b)
