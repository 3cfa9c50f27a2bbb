"""The FP64 fused step divides by its constants and by a cell's rho*dvc with
div_rn (x / b from the correctly rounded 1/b and two FMA residual
corrections, ark-cpp-mpi_amd/csrc/ark_device.hpp).  Its bitwise parity with
the reference rests on div_rn returning the IEEE quotient; this checks the
formula against the host's IEEE division (tests/native/div_rn_check.c)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_rn_is_ieee_division(tmp_path):
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-pthread", "-o", exe,
                    os.path.join(HERE, "native", "div_rn_check.c"), "-lm"], check=True)
    r = subprocess.run([exe, "40000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
