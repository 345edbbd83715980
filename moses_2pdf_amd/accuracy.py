"""Error of a reduced-precision run against the FP64 run of the same scenario.

Final flow depth h = hs + hf over the interior cells (scaled units, x10 of metres):
  l1    = sum |h - h64| / sum |h64|
  linf  = max |h - h64| / max |h64|
Runout extent: the set of cells with h above a threshold (default 0.01 m); reported as the
two cell counts, their relative difference and the intersection-over-union of the sets.
"""
from __future__ import annotations

import numpy as np


def flow_depth(u: np.ndarray) -> np.ndarray:
    u = np.asarray(u, np.float64)
    return u[0, 3:-3, 3:-3] + u[3, 3:-3, 3:-3]


def depth_error(u: np.ndarray, u64: np.ndarray, runout_m: float = 0.01) -> dict:
    h, h64 = flow_depth(u), flow_depth(u64)
    d = np.abs(h - h64)
    s64 = float(np.abs(h64).sum())
    m64 = float(np.abs(h64).max())
    thr = runout_m * 10.0
    a, b = h > thr, h64 > thr
    union = int(np.count_nonzero(a | b))
    return {
        "l1": float(d.sum()) / s64 if s64 > 0 else 0.0,
        "linf": float(d.max()) / m64 if m64 > 0 else 0.0,
        "runout_cells": int(np.count_nonzero(a)),
        "runout_cells_fp64": int(np.count_nonzero(b)),
        "runout_rel": abs(int(np.count_nonzero(a)) - int(np.count_nonzero(b))) / max(1, int(np.count_nonzero(b))),
        "runout_iou": (int(np.count_nonzero(a & b)) / union) if union else 1.0,
    }


# ---------------------------------------------------------------------------------------
# FP32 against FP64 at equal simulated time, and the conditioning floor
# ---------------------------------------------------------------------------------------

def mid_step_time(sc, steps: int) -> float:
    """A time half-way through the FP64 run's step number `steps` (both precisions then
    take the same number of steps to reach it, the last one shortened, whatever the float
    rounding of dt)."""
    from .engine import Simulation
    with Simulation(sc) as a:
        a.step(steps - 1)
        t0 = a.time_steps()[0]
        a.step(1)
        return 0.5 * (t0 + a.time_steps()[0])


def run_to_time(sc, t_end: float, max_steps: int, precision: str = "fp64", perturb: float = 0.0,
                seed: int = 1, metric_form: int = 0):
    """(state, mass, steps) of `sc` stepped with m2p_run_to to the scaled time t_end.
    perturb > 0: the initial state (an inflow run, which starts dry: the hydrograph depths)
    carries that relative random perturbation first.  metric_form: m2p_params.metric_form."""
    import dataclasses
    from .engine import Simulation
    rng = np.random.default_rng(seed)
    if perturb and sc.mode == 1:
        sc = dataclasses.replace(
            sc, inflow_h=sc.inflow_h * (1 + perturb * rng.standard_normal(sc.inflow_h.shape)))
    with Simulation(sc, precision=precision, metric_form=metric_form) as s:
        if perturb and sc.mode == 0:
            u = s.state()
            s.set_state(u * (1 + perturb * rng.standard_normal(u.shape)))
        n = s.run_to(t_end, max_steps)
        if s.time_steps()[0] != t_end:
            raise RuntimeError(f"run_to stopped at {s.time_steps()[0]} before {t_end}")
        return s.state(), s.mass(), n


def mass_rel(m, ref) -> float:
    return max(abs(x - y) / abs(y) for x, y in zip(m, ref))


def fp32_vs_fp64(sc, steps: int, floor: bool = True) -> dict:
    """The FP32 run against the FP64 run at a time half-way through FP64 step `steps`:
    L1/Linf of the final flow depth, runout IoU, mass difference, and (floor) the L1 of an
    FP64 run perturbed by 1e-7 relative (a float rounding) against the FP64 run."""
    t = mid_step_time(sc, steps)
    u64, m64, n64 = run_to_time(sc, t, 2 * steps)
    u32, m32, n32 = run_to_time(sc, t, 20 * steps, "fp32")
    e = depth_error(u32, u64)
    out = {"steps": n64, "steps_fp32": n32, "t_end": t, "l1": e["l1"], "linf": e["linf"],
           "runout_iou": e["runout_iou"], "mass_rel": mass_rel(m32, m64),
           "finite": bool(np.isfinite(u32).all())}
    if floor:
        up, _, _ = run_to_time(sc, t, 20 * steps, perturb=1e-7)
        out["floor_l1"] = depth_error(up, u64)["l1"]
    return out
