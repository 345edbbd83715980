"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py load
this module.  It restates, in numpy, the input handling of RunKernels (the
``stof``/``stod`` parse precision, the x10 scaling, the hydrograph tables and the
``dire`` rule), and hands the result to ``liboracle.so``.

References (MoSES_2PDF tree): L/R = MoSES_2PDF_landslide/adNOC_2Pb_runKernels.cu,
I/R = MoSES_2PDF_inflow/adNOC_2Pb_runKernels.cu.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_pad_input.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_new.restype = ctypes.c_void_p
        L.orc_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp,
                              dp, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp, dp, ip, dp]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.restype = ctypes.c_int
        L.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, dp]
        L.orc_get.restype = ctypes.c_int
        L.orc_set_u.argtypes = [ctypes.c_void_p, dp]
        L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_faces.argtypes = [ctypes.c_void_p]
        L.orc_inject.argtypes = [ctypes.c_void_p]
        L.orc_stage_terms.argtypes = [ctypes.c_void_p]
        L.orc_nxd.argtypes = [ctypes.c_void_p]
        L.orc_nyd.argtypes = [ctypes.c_void_p]
        L.orc_threads.argtypes = [ctypes.c_int]
        L.orc_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def threads(n: int = 0) -> int:
    """Sets (n > 0) and returns the oracle's OpenMP thread count."""
    return lib().orc_threads(int(n))


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def f32(x):
    """std::stof: the value rounded to float, then widened (L/R:146-152, 221)."""
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def pad(raw: np.ndarray) -> np.ndarray:
    """dataInitialization's x10 scaling, clamp and ghost copy (L/R:216-301)."""
    ny, nx = raw.shape
    out = np.zeros((ny + 6) * (nx + 6), dtype=np.float64)
    r = np.ascontiguousarray(raw, dtype=np.float64)
    lib().orc_pad_input(nx, ny, _dp(r), _dp(out))
    return out.reshape(ny + 6, nx + 6)


def inflow_tables(sc):
    """Hydrograph and location tables (I/R:297-331, 394-428)."""
    nsamp, nloc = sc.inflow_h.shape
    t = sc.inflow_t.astype(np.float32)
    inflowTime = (t * np.float32(10)).astype(np.float64)        # stof(..)*10 in float
    table = np.zeros((3, nsamp, nloc), dtype=np.float64)
    table[0] = 10.0 * f32(sc.inflow_h)
    table[1] = f32(sc.inflow_u)
    table[2] = f32(sc.inflow_v)
    loc = f32(sc.inflow_loc)                                    # [nloc, 3]
    dire = np.zeros((nloc, 3), dtype=np.int32)
    for j in range(nloc):
        dire[j, 0] = int(loc[j, 0])
        if loc[j, 2] == 0 or loc[j, 2] == sc.ny:
            dire[j, 1], dire[j, 2] = -5, 0
        else:
            dire[j, 1], dire[j, 2] = 0, -5
    return (np.ascontiguousarray(table.reshape(-1)), np.ascontiguousarray(loc.reshape(-1)),
            np.ascontiguousarray(dire.reshape(-1)), inflowTime)


def output_times(total_sim: float, each_out: float) -> np.ndarray:
    """tf[] of kernelStep (L/R:382-387), scaled by 10."""
    n = int(total_sim / each_out)
    tf = np.zeros(n + 1, dtype=np.float64)
    for ii in range(1, n + 1):
        tf[ii] = ii * each_out * 10
    return tf


class Oracle:
    """One simulation on the CPU oracle."""

    def __init__(self, sc, outputs: bool = False, tf=None):
        L = lib()
        self.sc = sc
        topo = pad(f32(sc.topo))
        depth = pad(sc.depth) if (sc.mode == 0 and sc.depth is not None) else None
        phi = None
        if sc.phi is not None:  # per-cell fraction: ghost copy only, no scaling
            phi = np.ascontiguousarray(np.pad(np.maximum(sc.phi, 0.0), 3, mode="edge"))
        dx = float(np.float32(sc.cellsize)) * 10
        par = np.array([sc.delta0, sc.Cd, sc.N_R, sc.varTheta, sc.phiS0, sc.cfl, sc.rho_ratio,
                        sc.dt_floor(), sc.dt_max], dtype=np.float64)
        if tf is not None:  # explicit clipping times (tf[0] = 0, as output_times gives)
            tf = np.ascontiguousarray(tf, dtype=np.float64)
            ntf = len(tf)
        else:
            tf = output_times(sc.total_sim, sc.each_out) if outputs else np.zeros(1)
            ntf = len(tf) if outputs else 0
        if sc.mode == 1:
            table, loc, dire, itime = inflow_tables(sc)
            nsamp, nloc = sc.inflow_h.shape
        else:
            table = loc = itime = np.zeros(1)
            dire = np.zeros(3, dtype=np.int32)
            nsamp = nloc = 0
        self._keep = (topo, depth, phi, par, tf, table, loc, dire, itime)
        self.h = L.orc_new(sc.mode, sc.nx, sc.ny, dx, _dp(topo), _dp(depth), _dp(phi), _dp(par),
                           ntf, _dp(tf), nloc, nsamp, _dp(table), _dp(loc),
                           dire.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(itime))
        self.nxd = sc.nx + 6
        self.nyd = sc.ny + 6
        self.P = self.nxd * self.nyd

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_free(self.h)
                self.h = None
        except Exception:
            pass

    def step(self, n: int = 1) -> int:
        return lib().orc_step(self.h, n)

    def get(self, name: str) -> np.ndarray:
        planes = {"u": 7, "uone": 7, "uzero": 7, "svec": 2, "bfkt": 3, "ux": 7, "uy": 7,
                  "uE": 7, "uW": 7, "uN": 7, "uS": 7, "Hpx": 7, "Hpy": 7, "Ppx": 7,
                  "Ppy": 7, "PDx": 7, "PDy": 7, "s": 7}.get(name, 1)
        out = np.zeros(planes * self.P, dtype=np.float64)
        n = lib().orc_get(self.h, name.encode(), _dp(out))
        if n < 0:
            raise KeyError(name)
        return out.reshape(planes, self.nyd, self.nxd) if planes > 1 else out.reshape(self.nyd, self.nxd)

    def faces(self) -> None:
        """First half of a predictor (no update): fills ux, uy, uE.., Val, dmax."""
        lib().orc_faces(self.h)

    def stage_terms(self) -> None:
        """faces(), then the numerical and viscous fluxes and the cell sources of that
        predictor (no update): fills PDx, PDy, the face gradients, Ac, AcF and the loads."""
        lib().orc_stage_terms(self.h)

    def inject(self) -> None:
        lib().orc_inject(self.h)

    def set_u(self, u: np.ndarray) -> None:
        a = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        assert a.size == 7 * self.P
        lib().orc_set_u(self.h, _dp(a))

    def scalar(self, name: str) -> float:
        return lib().orc_scalar(self.h, name.encode())
