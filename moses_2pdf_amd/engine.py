"""ctypes binding of ``libmoses2pdf.so`` (the C ABI in include/moses2pdf.h).

``RunKernels`` mirrors the reference's host class (MoSES_2PDF_landslide/
adNOC_2Pb_runKernels.cuh:27-377): construct it from a par_List path, call ``run()``.
``Simulation`` is the in-memory interface used by the tests, the benchmark and the
serving loop.  There is no CPU fallback: without the HIP library every call raises.
"""
from __future__ import annotations

import ctypes
import os
import time
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmoses2pdf.so")
_LIB = None

M2P_LANDSLIDE = 0
M2P_INFLOW = 1
M2P_FP64 = 0
M2P_FP32 = 1
PRECISIONS = {"fp64": M2P_FP64, "fp32": M2P_FP32}

EXPORTS = ["m2p_default_params", "m2p_open", "m2p_open_grid", "m2p_run", "m2p_step", "m2p_run_to", "m2p_sync",
           "m2p_get_state", "m2p_set_state", "m2p_dims", "m2p_mass", "m2p_scalars",
           "m2p_get_field", "m2p_profile", "m2p_profile_get", "m2p_last_error", "m2p_close",
           "m2p_set_stop", "m2p_status", "m2p_footprint", "m2p_checkpoint", "m2p_restore",
           "m2p_run_steps", "m2p_set_dry_skip", "m2p_nonfinite", "m2p_mass_noalloc",
           "m2p_batch_create", "m2p_batch_add", "m2p_batch_remove", "m2p_batch_size", "m2p_batch_status",
           "m2p_batch_step", "m2p_batch_sync", "m2p_batch_tune", "m2p_batch_set", "m2p_batch_get",
           "m2p_batch_last_error", "m2p_batch_destroy"]


class M2PError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("total_sim", "each_out", "delta0", "Cd", "N_R", "varTheta", "phiS0", "cfl",
                 "rho_ratio", "dt_min", "dt_max")] + [("precision", ctypes.c_int),
                                                       ("dry_skip", ctypes.c_int),
                                                       ("metric_form", ctypes.c_int)]


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise M2PError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    L.m2p_default_params.argtypes = [ctypes.c_int, ctypes.POINTER(Params)]
    L.m2p_default_params.restype = None
    L.m2p_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.m2p_open_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp,
                                dp, ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                dp, dp, ctypes.POINTER(vp)]
    L.m2p_run.argtypes = [vp, ctypes.c_char_p, dp]
    L.m2p_step.argtypes = [vp, ctypes.c_int]
    L.m2p_run_to.argtypes = [vp, ctypes.c_double, ctypes.c_int, ip]
    L.m2p_sync.argtypes = [vp]
    L.m2p_get_state.argtypes = [vp, dp, dp, ip]
    L.m2p_set_state.argtypes = [vp, dp]
    L.m2p_dims.argtypes = [vp, ip, ip, ip, ip]
    L.m2p_mass.argtypes = [vp, dp, dp]
    L.m2p_scalars.argtypes = [vp, dp, dp, ip]
    L.m2p_get_field.argtypes = [vp, ctypes.c_char_p, dp]
    L.m2p_profile.argtypes = [vp, ctypes.c_int]
    L.m2p_profile_get.argtypes = [vp, ctypes.c_char_p, dp, ip]
    L.m2p_last_error.argtypes = [vp]
    L.m2p_last_error.restype = ctypes.c_char_p
    L.m2p_close.argtypes = [vp]
    L.m2p_close.restype = None
    llp = ctypes.POINTER(ctypes.c_longlong)
    L.m2p_set_stop.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.m2p_set_dry_skip.argtypes = [vp, ctypes.c_int]
    L.m2p_checkpoint.argtypes = [vp, ctypes.c_char_p]
    L.m2p_restore.argtypes = [vp, ctypes.c_char_p]
    L.m2p_run_steps.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ip, dp]
    L.m2p_status.argtypes = [vp, dp, ip, ip]
    L.m2p_nonfinite.argtypes = [vp, ip, ip]
    L.m2p_mass_noalloc.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp, vp]
    L.m2p_footprint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, llp, llp]
    L.m2p_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.m2p_batch_add.argtypes = [vp, vp]
    L.m2p_batch_remove.argtypes = [vp, vp]
    L.m2p_batch_size.argtypes = [vp]
    L.m2p_batch_status.argtypes = [vp, ctypes.c_int, dp, ip, ip, ip, ip]
    L.m2p_batch_step.argtypes = [vp, ctypes.c_int]
    L.m2p_batch_sync.argtypes = [vp]
    L.m2p_batch_tune.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.m2p_batch_set.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.m2p_batch_get.argtypes = [vp, ctypes.c_char_p]
    L.m2p_batch_last_error.argtypes = [vp]
    L.m2p_batch_last_error.restype = ctypes.c_char_p
    L.m2p_batch_destroy.argtypes = [vp]
    L.m2p_batch_destroy.restype = None
    _LIB = L
    return L


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def default_params(mode: int) -> Params:
    p = Params()
    lib().m2p_default_params(mode, ctypes.byref(p))
    return p


class Simulation:
    """One scenario resident on the GPU.  ``precision`` is "fp64" (the reference's
    arithmetic) or "fp32" (FP32-bulk: float state and per-step arithmetic, FP64 metric
    setup, CFL max, dt control and mass totals).  ``dry_skip`` leaves out tiles whose
    inputs are all zero (the results are the same; see m2p_params.dry_skip).
    ``metric_form`` 1 makes the step load the stored metric planes even where the set-up
    check allows forming them (m2p_params.metric_form)."""

    def __init__(self, sc=None, *, par_list: Optional[str] = None, mode: Optional[int] = None,
                 overlay: Optional[str] = None, precision: str = "fp64", dry_skip: bool = True,
                 metric_form: int = 0):
        L = lib()
        self._h = ctypes.c_void_p()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self._dry_skip = bool(dry_skip)
        self._metric_form = int(metric_form)
        if par_list is not None:
            if precision != "fp64":
                overlay = ((overlay + ";") if overlay else "") + "precision=" + precision
            if not dry_skip:
                overlay = ((overlay + ";") if overlay else "") + "dry_skip=0"
            if self._metric_form != 0:
                overlay = ((overlay + ";") if overlay else "") + f"metric_form={self._metric_form}"
            rc = L.m2p_open(par_list.encode(), int(mode), overlay.encode() if overlay else None,
                            ctypes.byref(self._h))
        else:
            rc = self._open_scenario(L, sc)
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise M2PError(f"m2p_open failed ({rc}): {msg}")
        nx, ny, nxd, nyd = (ctypes.c_int() for _ in range(4))
        L.m2p_dims(self._h, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nxd), ctypes.byref(nyd))
        self.nx, self.ny, self.nxd, self.nyd = nx.value, ny.value, nxd.value, nyd.value
        self.P = self.nxd * self.nyd

    def _open_scenario(self, L, sc) -> int:
        p = default_params(sc.mode)
        p.total_sim, p.each_out = float(sc.total_sim), float(sc.each_out)
        p.delta0, p.Cd, p.N_R, p.varTheta, p.phiS0 = sc.delta0, sc.Cd, sc.N_R, sc.varTheta, sc.phiS0
        p.cfl, p.rho_ratio, p.dt_min, p.dt_max = sc.cfl, sc.rho_ratio, sc.dt_floor(), sc.dt_max
        p.precision = PRECISIONS[self.precision]
        p.dry_skip = 1 if self._dry_skip else 0
        p.metric_form = self._metric_form
        f32 = lambda a: None if a is None else np.ascontiguousarray(
            np.asarray(a, np.float64).astype(np.float32).astype(np.float64))
        topo = f32(sc.topo)
        cell = float(np.float32(sc.cellsize))
        depth = _f64(sc.depth) if sc.mode == 0 else None
        phi = _f64(sc.phi) if sc.mode == 0 else None
        if sc.mode == 1:
            t, h, u, v, loc = (f32(sc.inflow_t), f32(sc.inflow_h), f32(sc.inflow_u),
                               f32(sc.inflow_v), f32(sc.inflow_loc))
            nsamp, nloc = sc.inflow_h.shape
        else:
            t = h = u = v = loc = None
            nsamp = nloc = 0
        self._keep = (topo, depth, phi, t, h, u, v, loc)
        return L.m2p_open_grid(sc.mode, sc.nx, sc.ny, cell, _dp(topo), _dp(depth), _dp(phi),
                               ctypes.byref(p), nsamp, nloc, _dp(t), _dp(h), _dp(u), _dp(v),
                               _dp(loc), ctypes.byref(self._h))

    def _check(self, rc: int, what: str):
        if rc < 0:
            raise M2PError(f"{what} failed ({rc}): {self.last_error()}")
        return rc

    def last_error(self) -> str:
        if not self._h:
            return ""
        m = lib().m2p_last_error(self._h)
        return m.decode() if m else ""

    def step(self, n: int = 1) -> None:
        self._check(lib().m2p_step(self._h, int(n)), "m2p_step")

    def run_to(self, t_end: float, max_steps: int = 1 << 30) -> int:
        """Steps until the scaled simulated time reaches ``t_end`` (the last step shortened
        to end on it) or ``max_steps`` steps are done; returns the steps taken."""
        n = ctypes.c_int()
        self._check(lib().m2p_run_to(self._h, float(t_end), int(max_steps), ctypes.byref(n)),
                    "m2p_run_to")
        return n.value

    def sync(self) -> None:
        self._check(lib().m2p_sync(self._h), "m2p_sync")

    def set_stop(self, t_end: float = -1.0, vstop: float = -1.0, max_steps: int = -1) -> None:
        """Halting rules counted from now (see m2p_set_stop)."""
        self._check(lib().m2p_set_stop(self._h, float(t_end), float(vstop), int(max_steps)),
                    "m2p_set_stop")

    def set_dry_skip(self, on: bool) -> None:
        """Dry-tile skipping on or off from the next step (see m2p_set_dry_skip)."""
        self._check(lib().m2p_set_dry_skip(self._h, 1 if on else 0), "m2p_set_dry_skip")

    def status(self):
        """(scaled time, steps done, halted)."""
        t, nt, h = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().m2p_status(self._h, ctypes.byref(t), ctypes.byref(nt), ctypes.byref(h)),
                    "m2p_status")
        return t.value, nt.value, bool(h.value)

    def nonfinite(self):
        """(flag, first_nt): whether a CFL pass has met a non-finite state, and the step
        count of the first such state (-1 if none); see m2p_nonfinite."""
        f, nt = ctypes.c_int(), ctypes.c_int()
        self._check(lib().m2p_nonfinite(self._h, ctypes.byref(f), ctypes.byref(nt)), "m2p_nonfinite")
        return bool(f.value), nt.value

    def state(self) -> np.ndarray:
        u = np.empty(7 * self.P, dtype=np.float64)
        self._check(lib().m2p_get_state(self._h, _dp(u), None, None), "m2p_get_state")
        return u.reshape(7, self.nyd, self.nxd)

    def time_steps(self):
        t = ctypes.c_double()
        nt = ctypes.c_int()
        self._check(lib().m2p_get_state(self._h, None, ctypes.byref(t), ctypes.byref(nt)), "m2p_get_state")
        return t.value, nt.value

    def set_state(self, u: np.ndarray) -> None:
        a = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        assert a.size == 7 * self.P
        self._check(lib().m2p_set_state(self._h, _dp(a)), "m2p_set_state")

    def mass(self):
        s, f = ctypes.c_double(), ctypes.c_double()
        self._check(lib().m2p_mass(self._h, ctypes.byref(s), ctypes.byref(f)), "m2p_mass")
        return s.value, f.value

    def scalars(self):
        dt, dmax, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(lib().m2p_scalars(self._h, ctypes.byref(dt), ctypes.byref(dmax), ctypes.byref(it)),
                    "m2p_scalars")
        return dt.value, dmax.value, it.value

    def field(self, name: str) -> np.ndarray:
        planes = {"svec": 2, "uone45": 2}.get(name, 1)
        out = np.empty(planes * self.P, dtype=np.float64)
        self._check(lib().m2p_get_field(self._h, name.encode(), _dp(out)), "m2p_get_field")
        return out.reshape(planes, self.nyd, self.nxd) if planes > 1 else out.reshape(self.nyd, self.nxd)

    def profile(self, enable: bool) -> None:
        self._check(lib().m2p_profile(self._h, int(enable)), "m2p_profile")

    def profile_get(self, kernel: str):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._check(lib().m2p_profile_get(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)),
                    "m2p_profile_get")
        return ms.value, n.value

    def run(self, result_dir: str) -> float:
        w = ctypes.c_double()
        self._check(lib().m2p_run(self._h, result_dir.encode(), ctypes.byref(w)), "m2p_run")
        return w.value

    def run_steps(self, result_dir: str, max_steps: int) -> bool:
        """m2p_run for at most max_steps steps; True once the last output is written."""
        fin, w = ctypes.c_int(), ctypes.c_double()
        self._check(lib().m2p_run_steps(self._h, result_dir.encode(), int(max_steps), ctypes.byref(fin),
                                        ctypes.byref(w)), "m2p_run_steps")
        return bool(fin.value)

    def checkpoint(self, path: str) -> None:
        self._check(lib().m2p_checkpoint(self._h, path.encode()), "m2p_checkpoint")

    def restore(self, path: str) -> None:
        self._check(lib().m2p_restore(self._h, path.encode()), "m2p_restore")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().m2p_close(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def footprint(mode: int, nx: int, ny: int, precision: str = "fp64", nsamp: int = 0,
              nloc: int = 0):
    """(steady, peak) device bytes of a scenario of that shape."""
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    rc = lib().m2p_footprint(int(mode), int(nx), int(ny), PRECISIONS[precision], int(nsamp),
                             int(nloc), ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise M2PError(f"m2p_footprint failed ({rc})")
    return a.value, b.value


class Batch:
    """Many scenarios of one precision stepped together (m2p_batch_*)."""

    def __init__(self, precision: str = "fp64"):
        self._h = ctypes.c_void_p()
        self.precision = precision
        self.members = []
        rc = lib().m2p_batch_create(PRECISIONS[precision], ctypes.byref(self._h))
        if rc:
            msg = self.last_error()
            self.close()
            raise M2PError(f"m2p_batch_create failed ({rc}): {msg}")

    def last_error(self) -> str:
        if not self._h:
            return ""
        m = lib().m2p_batch_last_error(self._h)
        return m.decode() if m else ""

    def _check(self, rc: int, what: str):
        if rc < 0:
            raise M2PError(f"{what} failed ({rc}): {self.last_error()}")
        return rc

    def add(self, sim: "Simulation") -> int:
        k = self._check(lib().m2p_batch_add(self._h, sim._h), "m2p_batch_add")
        self.members.append(sim)
        return k

    def remove(self, sim: "Simulation") -> None:
        self._check(lib().m2p_batch_remove(self._h, sim._h), "m2p_batch_remove")
        self.members.remove(sim)

    def __len__(self) -> int:
        return self._check(lib().m2p_batch_size(self._h), "m2p_batch_size")

    def status(self):
        """Per member, in order of addition (self.members): (scaled time, steps done, halted,
        non-finite, first non-finite step or -1), with one device-to-host copy for the batch
        (m2p_batch_status)."""
        n = len(self.members)
        t = np.zeros(max(n, 1), np.float64)
        nt, h, bad, bad_nt = (np.zeros(max(n, 1), np.int32) for _ in range(4))
        self._check(lib().m2p_batch_status(self._h, n, _dp(t), _ip(nt), _ip(h), _ip(bad), _ip(bad_nt)),
                    "m2p_batch_status")
        return [(float(t[k]), int(nt[k]), bool(h[k]), bool(bad[k]), int(bad_nt[k])) for k in range(n)]

    def step(self, n: int = 1) -> None:
        self._check(lib().m2p_batch_step(self._h, int(n)), "m2p_batch_step")

    def sync(self) -> None:
        self._check(lib().m2p_batch_sync(self._h), "m2p_batch_sync")

    def tune(self, trial_steps: int = 3) -> str:
        buf = ctypes.create_string_buffer(8192)
        self._check(lib().m2p_batch_tune(self._h, int(trial_steps), buf, len(buf)), "m2p_batch_tune")
        return buf.value.decode()

    def set(self, key: str, value: int) -> None:
        self._check(lib().m2p_batch_set(self._h, key.encode(), int(value)), "m2p_batch_set")

    def get(self, key: str) -> int:
        return self._check(lib().m2p_batch_get(self._h, key.encode()), "m2p_batch_get")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().m2p_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class RunKernels:
    """Mirror of the reference's RunKernels (adNOC_2Pb_runKernels.cuh:27-377).

    ``RunKernels(fname, mode)`` parses ``fname`` (a par_List) and its data files;
    ``run()`` writes ``./result2Pb/`` and returns the elapsed time in clock ticks of
    ``time.process_time`` scaled like ``clock()`` (CLOCKS_PER_SEC = 1e6)."""

    def __init__(self, fname: str = "./par_List", mode: int = M2P_LANDSLIDE,
                 overlay: Optional[str] = None):
        self.sim = Simulation(par_list=fname, mode=mode, overlay=overlay)

    def run(self, result_dir: str = "./result2Pb/") -> int:
        t0 = time.process_time()
        self.sim.run(result_dir)
        return int((time.process_time() - t0) * 1e6)
