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
