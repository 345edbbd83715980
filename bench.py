#!/usr/bin/env python3
"""Throughput benchmark of the adNOC_2Pb time-step loop on one MI355X.

Workload: landslide mode, plain FP64, 2048x2048 cells at 20 m on a seeded diamond-square
DEM (Hurst 0.8, 1500 m relief, 3-cell box filter) with three seeded paraboloid releases
(1e5..5e6 m^3, solid fraction 0.4..0.7).  One step = one full time step (CFL max
reduction, predictor, corrector) over every interior cell; the state stays resident in
HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2048] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic HBM bytes per interior cell for each pass of one step (FP64): the state
# planes a pass must read and write plus the topography, with the metric recomputable
# from the topography (so the stored-metric reads are not counted).
PLANES = {
    "cfl": 7 + 1,                 # read u^n
    "stage0": (7 + 7 + 2) + 1,    # read u^n, write u^{n+1/2} and 2 limiter planes
    "stage1": (7 + 6 + 2 + 7) + 1,  # read u^{n+1/2}, u^n, limiter; write u^{n+1}
}
ELEM = {"fp64": 8, "fp32": 4}
ALG_BYTES = {k: v * 8 for k, v in PLANES.items()}


def traffic_from_profiles(kernel: str, n: int):
    """HBM bytes per launch of `kernel` at grid n, from the committed rocprofv3 PMC summary
    (profiles/hbm_traffic_<n>.txt lines 'kernel: bytes_per_launch'), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{n}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def cpu_baseline(sc, steps: int):
    """The FP64 oracle (serial C++ restatement of the reference kernels) on the same
    scenario, timed around `steps` steps only."""
    from oracle.oracle import Oracle
    o = Oracle(sc)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": sc.nx * sc.ny * steps / dt, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} full step(s) of the same {sc.nx}x{sc.ny} landslide workload, "
                      f"serial oracle (-O2 -ffp-contract=off), step loop only, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--precision", choices=sorted(ELEM), default="fp64",
                    help="fp64: the reference's arithmetic (the headline); fp32: FP32-bulk")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after the timed steps, copy the 7 state planes once device to device "
                         "(a plain copy kernel of known size, for HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (one MI355X per box); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    from moses_2pdf_amd import Simulation
    from moses_2pdf_amd import scenarios as S

    sc = S.landslide_fractal(n=args.n, seed=2)
    print(f"# workload ready: {sc.nx}x{sc.ny}", file=sys.stderr, flush=True)
    sim = Simulation(sc, precision=args.precision)
    alg_bytes = {k: v * ELEM[args.precision] for k, v in PLANES.items()}
    sim.step(args.warmup)
    sim.sync()
    print(f"# warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    sim.profile(True)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    kern = {k: sim.profile_get(k) for k in ("ghost", "cfl", "dt", "stage0", "stage1")}
    sim.profile(False)
    elapsed = t1 - t0
    cells = sc.nx * sc.ny
    value = cells * args.steps / elapsed
    print(f"# timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    u = sim.state()
    import numpy as np
    if args.copy_calib:
        sim.set_state(u)   # H2D of 7 planes, then one D2D copy of 7*nxd*nyd doubles
        sim.sync()
        print(f"# copy-calib: D2D copy of {7 * sim.P * 8} bytes", file=sys.stderr, flush=True)
    ok = bool(np.isfinite(u).all())
    ms, mf = sim.mass()

    totals = {k: v[0] * v[1] for k, v in kern.items()}
    dom = max(("cfl", "stage0", "stage1"), key=lambda k: totals[k])
    avg_ms = kern[dom][0]
    alg = alg_bytes[dom] * cells
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(dom, args.n) if args.precision == "fp64" else None
    step_alg = sum(alg_bytes.values()) * cells
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (f64 dt, CFL max, mass)",
        "data": "synthetic",
        "config": {"workload": f"landslide {args.precision.upper()}, {sc.nx}x{sc.ny} cells at 20 m, seeded "
                               "diamond-square DEM (H 0.8, 1500 m relief, 3-cell box), "
                               "3 paraboloid releases, per-cell phi 0.4-0.7",
                   "grid": [sc.nx, sc.ny], "cfl": sc.cfl, "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_cell": alg_bytes[dom]},
        "step_hbm_alg_GBs": step_alg * args.steps / elapsed / 1e9,
        "kernels_ms": {k: v[0] for k, v in kern.items()},
        "state_finite": ok,
        "mass": [ms, mf],
    }
    if not args.no_cpu_baseline:
        print("# cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        del sim
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
