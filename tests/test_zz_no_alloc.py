"""ark_diagnostics_noalloc against ark_diagnostics, through the C-ABI.

The caller's device buffers are allocated with the HIP runtime the library
itself links (libamdhip64.so.7, already loaded by it), as a caller that owns
its fields would; the entry point under test allocates nothing.  Two calls in
a row on one GPU, the second on a smaller grid, share the library's scratch
array; each must give the context's sums bit for bit.
"""
import ctypes

import numpy as np
import pytest

import arkamd

pytestmark = pytest.mark.gpu

CELLS = ("ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon")
KEYS = ("ke_reference", "ke_physical", "mass", "max_abs_u", "nonfinite")
H2D, D2H = 1, 2


class DeviceBuffers:
    """hipMalloc'd copies of host arrays (the caller's side)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0   # synchronises
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc(solver, host_cells, dev):
    ptrs = (ctypes.c_void_p * 5)(*[dev.put(host_cells[n]) for n in CELLS])
    out = dev.alloc(5 * 8)
    st = arkamd.lib().ark_diagnostics_noalloc(ctypes.byref(solver.params), ptrs, out, None)
    assert st == arkamd.ARK_OK, st
    return dev.get(out, 5)


def bits(x):
    return np.float64(x).view(np.int64)


@pytest.mark.parametrize("nan", [False, True])
def test_noalloc_diagnostics_equal_context(nan):
    dev = DeviceBuffers()
    try:
        for N in ((40, 36, 32), (20, 18, 16)):       # the second call smaller
            with arkamd.Solver(*N, scheme=arkamd.SCHEME_CORRECTED) as s:
                s.init_tgv()
                s.step(3)
                if nan:
                    u = s.download("u2nCon")
                    u[4, 5, 6] = np.nan
                    s.upload("u2nCon", u)
                want = s.diagnostics(allow_nonfinite=True)
                host = {n: s.download(n) for n in CELLS}
                got = noalloc(s, host, dev)
            for q, k in enumerate(KEYS):
                if k == "nonfinite":
                    assert got[q] == want[k] == (1 if nan else 0), (N, k, got[q], want[k])
                elif nan and k in ("ke_reference", "ke_physical"):   # sums over the NaN cell
                    assert np.isnan(got[q]) and np.isnan(want[k]), (N, k, got[q], want[k])
                else:
                    assert bits(got[q]) == bits(want[k]), (N, k, got[q], want[k])
    finally:
        dev.free()


def test_noalloc_refuses_oversized_grid():
    p = arkamd.Params()
    p.N1 = p.N2 = p.N3 = 1400                   # 1402^3 > 2^31 elements
    p.dx1 = p.dx2 = p.dx3 = 1.0
    fake = (ctypes.c_void_p * 5)(*([16] * 5))   # never dereferenced: refused first
    st = arkamd.lib().ark_diagnostics_noalloc(ctypes.byref(p), fake, ctypes.c_void_p(16), None)
    assert st == arkamd.ARK_ERR_UNSUPPORTED
