"""m2p_mass_noalloc (the allocation-free variant of m2p_mass) against m2p_mass, through the
C ABI, on two calls in a row, the second on a smaller grid: the totals must be equal bit
for bit, in both precisions.  Device buffers come from the HIP runtime the library itself
uses, so the call sees them as the caller's own memory."""
import ctypes
import struct

import numpy as np
import pytest

from moses_2pdf_amd import Simulation
from moses_2pdf_amd import engine as E
from moses_2pdf_amd import scenarios as S

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
    return p


def _case(sc, steps, precision):
    with Simulation(sc, precision=precision) as s:
        s.step(steps)
        u = s.state()
        det = np.ascontiguousarray(s.field("Detmin"))
        want = s.mass()
    u = np.ascontiguousarray(u if precision == "fp64" else u.astype(np.float32))
    return sc.nx, sc.ny, u, det, want


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mass_noalloc_equals_mass(precision):
    hip = _hip()
    L = E.lib()
    prec = E.PRECISIONS[precision]
    cases = [_case(S.landslide_fractal(n=200, seed=5), 30, precision),
             _case(S.landslide_plane(n=64), 12, precision)]        # the second one smaller
    out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), 16) == 0
    held = [out]
    try:
        for nx, ny, u, det, want in cases:
            du, dd = _dev(hip, u), _dev(hip, det)
            held += [du, dd]
            assert L.m2p_mass_noalloc(du, prec, dd, nx, ny, out, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            got = (ctypes.c_double * 2)()
            assert hip.hipMemcpy(ctypes.cast(got, ctypes.c_void_p), out, 16, D2H) == 0
            print(precision, nx, ny, tuple(got), want, flush=True)
            assert struct.pack("dd", *got) == struct.pack("dd", *want)
        # rejected before any device work, with a message
        assert L.m2p_mass_noalloc(held[1], prec, held[2], 0, 64, out, None) == -1
        assert b"grid" in L.m2p_last_error(None)
        assert L.m2p_mass_noalloc(held[1], 7, held[2], 64, 64, out, None) == -1
        assert b"precision" in L.m2p_last_error(None)
    finally:
        for p in held:
            hip.hipFree(p)
