"""CPU checks of the C-ABI library: it loads and exports every symbol that
include/ark/ark.h declares (no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import REPO
import arkamd


def declared_symbols():
    src = open(os.path.join(REPO, "include", "ark", "ark.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ark_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(arkamd.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(arkamd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_params_struct_layout():
    # ark_params: 3 ints, 8 doubles, int[3], 2 ints -> offsets as the C compiler lays them out
    P = arkamd.Params
    assert P.dx1.offset == 16 and P.t0.offset == 72 and P.periodic.offset == 80
    assert P.precision.offset == 92 and ctypes.sizeof(P) == 104


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(arkamd, "_lib", None)
    monkeypatch.setattr(arkamd, "LIB_PATH", "/nonexistent/libark_amd.so")
    with pytest.raises(RuntimeError, match="not built"):
        arkamd.lib()


def test_cpp_host_layer_built():
    """The C++ host interface is in the library and its programs link against
    it (no GPU call: the programs are only asked for their usage)."""
    import subprocess
    lib = arkamd.LIB_PATH
    syms = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for name in ("ark::Solver::Solver", "ark::Arr3d::bind", "ark::Arr2d::pack",
                 "ark::BoundaryConditions::alongX1", "ark::BoundaryConditions::alongX3"):
        assert name in syms, name
    bindir = os.path.join(os.path.dirname(lib), "..", "bin")
    for prog in ("ark_driver", "ark_host_check"):
        res = subprocess.run([os.path.join(bindir, prog)], capture_output=True, text=True)
        assert res.returncode == 2 and "usage" in res.stderr


def test_create_refuses_boxes_beyond_32_bit_offsets():
    """The fused kernels address a field with 32-bit byte offsets, so a box
    with more than 2^29 elements per field is refused at creation (before
    any device work, so this runs without a GPU)."""
    for N in ((1024, 1024, 1024), (810, 810, 810), (4096, 512, 256)):
        with pytest.raises(arkamd.ArkError) as e:
            arkamd.Solver(*N)
        assert e.value.status == arkamd.ARK_ERR_UNSUPPORTED
        assert "32 bits" in str(e.value)
