#!/usr/bin/env python3
"""Throughput benchmark of the adNOC_2Pb time-step loop on one MI355X.

Headline workload: landslide mode, plain FP64, 2048x2048 cells at 20 m on a seeded
diamond-square DEM (Hurst 0.8, 1500 m relief, 3-cell box filter) with three seeded
paraboloid releases (1e5..5e6 m^3, solid fraction 0.4..0.7).  One step = one full time
step (CFL max reduction, predictor, corrector) over every interior cell; the state stays
resident in HBM.  The same JSON line carries, beside it, the batched-ensemble workload
(256 scenarios, 3000 steps: batched FP32 and FP64 against the plain FP64 path, and the
FP32 error), the serving workload (64 Poisson requests, continuous batching against the
plain path, p50/p99 latency) and the CPU baseline (the FP64 oracle in fixed windows).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2048] [--no-cpu-baseline]
                    [--no-ensemble] [--no-serving]
    python bench.py --workload ensemble|serve|accuracy     (one workload on its own line)
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


def traffic_from_profiles(kernel: str, n: int, precision: str = "fp64"):
    """HBM bytes per launch of `kernel` at grid n, from the committed rocprofv3 PMC summary
    (profiles/hbm_traffic[_fp32]_<n>.txt lines 'kernel: bytes_per_launch'), or None."""
    tag = "" if precision == "fp64" else f"{precision}_"
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}{n}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def valu_from_profiles(kernel: str, n: int, precision: str = "fp64"):
    """(VALU issue share, wait share, VALU busy share or None) of `kernel` at grid n from the
    committed rocprofv3 PMC summary (profiles/valu_counters_<n>.txt lines '<precision>
    <kernel>: ... VALU issue share A%, wait share B%, VALU busy C%'), or None."""
    path = os.path.join(ROOT, "profiles", f"valu_counters_{n}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"{precision} {kernel}:") and "issue share" in line:
            try:
                a = float(line.split("VALU issue share")[1].split("%")[0])
                b = float(line.split("wait share")[1].split("%")[0])
                c = (float(line.split("VALU busy")[1].split("%")[0]) / 100.0
                     if "VALU busy" in line else None)
                return a / 100.0, b / 100.0, c
            except (IndexError, ValueError):
                return None
    return None


def _cpu_info() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def _oracle_rate(sc, steps: int) -> tuple:
    """Cell-updates/s of the FP64 oracle on `sc`, timed around `steps` steps after one
    untimed step (first touch of its work arrays); (rate, seconds)."""
    from oracle import oracle as O
    o = O.Oracle(sc)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return sc.nx * sc.ny * steps / dt, dt


def cpu_baseline(sc_c2, c2_steps: int = 20, c1_steps: int = 500, c3_steps: int = 20,
                 serial_steps: int = 20) -> dict:
    """The FP64 oracle (the C++ restatement of the reference kernels, -O2 -ffp-contract=off,
    OpenMP over the host's threads; its results do not depend on the thread count) timed
    around the step loop only, in fixed windows of the three plain configurations: the
    2048^2 landslide (the headline workload, `value`), the 256^2 landslide on the 35-degree
    plane (CFL 0.4) and the 1024 x 4096 inflow valley; and the 256^2 case on one thread."""
    from moses_2pdf_amd import scenarios as S
    from oracle import oracle as O
    nthr = O.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    v2, d2 = _oracle_rate(sc_c2, c2_steps)
    log(f"cpu baseline: 2048^2 landslide, {c2_steps} steps in {d2:.1f} s on {nthr} threads")
    c1 = S.landslide_plane(n=256, phi=0.6)
    c1.cfl, c1.rho_ratio = 0.4, 1000.0 / 2700.0
    v1, d1 = _oracle_rate(c1, c1_steps)
    log(f"cpu baseline: 256^2 plane, {c1_steps} steps in {d1:.1f} s")
    c3 = S.inflow_valley()
    v3, d3 = _oracle_rate(c3, c3_steps)
    log(f"cpu baseline: 1024x4096 inflow valley, {c3_steps} steps in {d3:.1f} s")
    O.threads(1)
    vs, ds = _oracle_rate(c1, serial_steps)
    O.threads(nthr)
    log(f"cpu baseline: 256^2 plane, {serial_steps} steps on 1 thread in {ds:.1f} s")
    out = {"value": v2, "unit": "cell-updates/s", "cores": nthr, "kind": "port",
           "sample": f"window of {c2_steps} full steps of the same {sc_c2.nx}x{sc_c2.ny} landslide "
                     f"workload after one untimed step, FP64 oracle (-O2 -ffp-contract=off, "
                     f"OpenMP) on {nthr} threads, step loop only, {d2:.1f} s"}
    out.update(_cpu_info())
    out["windows"] = {
        "landslide_2048": {"steps": c2_steps, "value": v2, "seconds": d2, "threads": nthr},
        "landslide_plane_256": {"steps": c1_steps, "of_steps": 2000, "value": v1, "seconds": d1,
                                "threads": nthr},
        "inflow_valley_1024x4096": {"steps": c3_steps, "of_steps": 10000, "value": v3, "seconds": d3,
                                    "threads": nthr},
        "landslide_plane_256_serial": {"steps": serial_steps, "value": vs, "seconds": ds, "threads": 1},
    }
    return out


def accuracy(args) -> int:
    """The FP32-bulk variant against FP64 at equal simulated time on five scenarios (the
    first 128^2, 256^2 and 1024^2 ensemble members and the reference's two datasets): L1
    of the final flow depth after a short horizon (30 steps; 10 for the inflow case, whose
    hydrograph sample switches on a time test) and after 1000 steps, the conditioning floor
    (L1 of an FP64 run perturbed by 1e-7 relative) and the runout IoU at 1000 steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from util import df004, hsiaolin   # the reference's datasets, kept as test fixtures
    from moses_2pdf_amd import scenarios as S
    from moses_2pdf_amd.accuracy import fp32_vs_fp64
    ms = S.ensemble_members(256)
    first = lambda n: S.ensemble_scenario([m for m in ms if m.n == n][0])
    cases = {"ensemble_member_128": lambda: first(128), "ensemble_member_256": lambda: first(256),
             "ensemble_member_1024": lambda: first(1024), "hsiaolin": hsiaolin, "df004": df004}
    rows = {}
    for name, make in cases.items():
        sc = make()
        short = fp32_vs_fp64(sc, 10 if sc.mode == 1 else 30, floor=False)
        long_ = fp32_vs_fp64(sc, 1000)
        rows[name] = {"short_steps": short["steps"], "short_l1": short["l1"],
                      "l1_1000": long_["l1"], "floor_l1_1000": long_["floor_l1"],
                      "runout_iou_1000": long_["runout_iou"], "mass_rel_1000": long_["mass_rel"],
                      "linf_1000": long_["linf"]}
        print(f"# {name}: {rows[name]}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "FP32 final-depth L1 vs FP64", "cases": rows}), flush=True)
    return 0


def ensemble_measure(args, dense_steps=None) -> dict:
    """The 256-member batched ensemble: batched FP32 and FP64 (autotuned, graph-captured)
    against the plain FP64 path (each member alone with m2p_step), all with dry-tile
    skipping (unless --dense), and the FP32 final-depth error against FP64 at equal time.
    dense_steps: also time every path that computes every tile, over the first dense_steps
    of the steps (a dense step's cost does not depend on the state, so the full run's cost
    is the per-step cost times the steps)."""
    from moses_2pdf_amd import ensemble as E
    import numpy as np
    t0 = time.perf_counter()
    ms, scs = E.population(args.members, seed=4)
    E.log(f"population of {len(scs)} built in {time.perf_counter() - t0:.1f} s")
    cells = sum(sc.nx * sc.ny for sc in scs)
    steps = args.ens_steps
    dry = not args.dense
    n = len(scs)
    per_h = lambda wall: n / wall * 3600.0
    plain = E.run_plain(scs, steps, "fp64", dry_skip=dry)
    ref = plain["out"]
    b64 = E.run_batch(scs, steps, "fp64", keep=False, dry_skip=dry)
    b32 = E.run_batch(scs, steps, "fp32", keep=False, dry_skip=dry)
    # FP32 error at equal time: each member stepped to its FP64 run's final time (members
    # whose FP64 run went non-finite are left out: the reference scheme itself fails there)
    keep = [k for k in range(n) if not plain["nonfinite"][k] and np.isfinite(ref[k][0]).all()]
    # (a member whose FP32 dt collapses to its floor cannot reach the FP64 time within 4x the
    # steps: it is counted as stalled, not compared)
    acc = E.run_batch([scs[k] for k in keep], 4 * steps, "fp32", tune_steps=0,
                      t_end=[ref[k][1] for k in keep])
    fin32 = [bool(np.isfinite(o[0]).all()) and not f for o, f in zip(acc["out"], acc["nonfinite"])]
    reached = [acc["out"][j][1] == ref[keep[j]][1] for j in range(len(keep))]
    both = [j for j in range(len(keep)) if fin32[j] and reached[j]]
    err, l1 = E.errors([acc["out"][j][0] for j in both], [ref[keep[j]][0] for j in both])
    err["members_nonfinite_fp64"] = n - len(keep)
    err["members_nonfinite_fp32"] = len(keep) - sum(fin32)
    err["members_stalled_fp32"] = sum(1 for j in range(len(keep)) if fin32[j] and not reached[j])
    ok = [keep[j] for j in both]
    fidx, fl1 = E.floor_l1([scs[k] for k in ok], steps, [ref[k] for k in ok], count=args.floor_members)
    err["floor_l1_median_subset"] = float(np.median(fl1))
    err["fp32_over_floor_median_subset"] = float(np.median(l1[fidx] / np.maximum(fl1, 1e-12)))
    err["frac_l1_le_2floor_plus_1pct_subset"] = float((l1[fidx] <= 2 * fl1 + 1e-2).mean())
    del ref, plain["out"]
    out = {
        "scenarios_per_hour": per_h(b32["wall_s"]),
        "members": n, "cells": cells, "steps": steps, "dry_skip": dry,
        "speedup_vs_plain_fp64": plain["wall_s"] / b32["wall_s"],
        "plain_fp64": {"scenarios_per_hour": per_h(plain["wall_s"]), "wall_s": plain["wall_s"],
                       "members_nonfinite": int(sum(plain["nonfinite"]))},
        "batched_fp64": {"scenarios_per_hour": per_h(b64["wall_s"]), "wall_s": b64["wall_s"],
                         "step_s": b64["step_s"], "tune": b64["tune_report"].strip().split("\n")[-1],
                         "members_nonfinite": int(sum(b64["nonfinite"]))},
        "batched_fp32": {"scenarios_per_hour": per_h(b32["wall_s"]), "wall_s": b32["wall_s"],
                         "step_s": b32["step_s"], "open_s": b32["open_s"], "tune_s": b32["tune_s"],
                         "tune": b32["tune_report"].strip().split("\n")[-1],
                         "cell_updates_per_s": cells * steps / b32["step_s"],
                         "members_nonfinite": int(sum(b32["nonfinite"]))},
        "fp32_error_vs_fp64": err,
    }
    if dense_steps and dry:
        scale = steps / dense_steps
        pd = E.run_plain(scs, dense_steps, "fp64", keep=False, dry_skip=False)
        bd = E.run_batch(scs, dense_steps, "fp32", keep=False, dry_skip=False)
        wp = pd["wall_s"] + pd["step_s"] * (scale - 1.0)
        wb = bd["wall_s"] + bd["step_s"] * (scale - 1.0)
        out["dense"] = {"timed_steps": dense_steps, "plain_fp64_scenarios_per_hour": per_h(wp),
                        "batched_fp32_scenarios_per_hour": per_h(wb), "speedup_vs_plain_fp64": wp / wb,
                        "note": f"every tile computed; stepping timed over {dense_steps} of the "
                                f"{steps} steps and scaled (a dense step's cost does not depend on "
                                "the state), set-up and tuning as measured"}
    return out


def ensemble(args) -> int:
    """The batched-ensemble workload on its own JSON line."""
    r = ensemble_measure(args)
    steps = r["steps"]
    out = {
        "metric": "ensemble scenarios/hour", "value": r["scenarios_per_hour"], "unit": "scenarios/h",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": r["batched_fp32"]["step_s"] / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 dt, CFL max, mass)",
        "data": "synthetic",
        "config": {"workload": ENSEMBLE_WORKLOAD.format(n=r["members"], steps=steps),
                   "members": r["members"], "cells": r["cells"], "steps": steps,
                   "dry_skip": r["dry_skip"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("scenarios_per_hour",)})
    print(json.dumps(out), flush=True)
    return 0


ENSEMBLE_WORKLOAD = ("batched ensemble, {n} landslides on seeded fractal DEMs (self-affine relief), "
                     "sides 128/256/512/1024 with weights 4:3:2:1, LHS basal friction 15-30 deg, "
                     "volume 1e5-5e6 m^3, phi 0.4-0.7, {steps} steps each, per-scenario dt")


def serve_measure(args) -> dict:
    """Continuous-batching FP32 serving against the plain FP64 one-at-a-time path on the
    same Poisson request stream."""
    from moses_2pdf_amd import serving as V
    t0 = time.perf_counter()
    reqs = V.make_requests(args.requests, rate=args.rate, seed=5)
    cells = sum(r.n * r.n for r in reqs)
    V_log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    V_log(f"{len(reqs)} requests built in {time.perf_counter() - t0:.1f} s "
          f"({sum(r.mode for r in reqs)} inflow)")
    dry = not args.dense
    srv = V.Server(precision="fp32", hbm_budget=args.hbm_budget * 1e9, tick_steps=args.tick,
                   max_steps=args.max_steps, dry_skip=dry)
    t1 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t1
    cb = V.summary(res, wall, args.max_steps)
    cb.update({"peak_active": srv.peak_active, "peak_device_GB": srv.peak_held / 1e9})
    V_log("continuous batching:", cb)
    t1 = time.perf_counter()
    res0 = V.serve_alone(reqs, precision="fp64", max_steps=args.max_steps, tick_steps=args.tick,
                         dry_skip=dry)
    wall0 = time.perf_counter() - t1
    plain = V.summary(res0, wall0, args.max_steps)
    V_log("plain fp64 one at a time:", plain)
    return {"runs_per_hour": cb["runs_per_hour"], "latency_p50_s": cb["latency_p50_s"],
            "latency_p99_s": cb["latency_p99_s"], "requests": len(reqs), "cells": cells,
            "dry_skip": dry, "continuous_batching_fp32": cb, "plain_fp64_one_at_a_time": plain,
            "speedup_runs_per_hour": cb["runs_per_hour"] / plain["runs_per_hour"]}


SERVE_WORKLOAD = ("serving: {k} hazard requests, Poisson {rate}/s, grid side lognormal (median 384, "
                  "sigma 0.5) clipped to 128-2048, landslide/inflow 70/30, stop at max flow speed "
                  "< 0.01 m/s after motion or {ms} steps, admission by HBM footprint ({gb} GB budget)")


def serve(args) -> int:
    """The serving workload on its own JSON line."""
    r = serve_measure(args)
    out = {
        "metric": "serving runs/hour", "value": r["runs_per_hour"], "unit": "runs/h", "n_gpus": 1,
        "steps": args.max_steps, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 dt, CFL max, mass)",
        "data": "synthetic",
        "config": {"workload": SERVE_WORKLOAD.format(k=r["requests"], rate=args.rate, ms=args.max_steps,
                                                     gb=args.hbm_budget),
                   "requests": r["requests"], "cells": r["cells"], "tick_steps": args.tick,
                   "dry_skip": r["dry_skip"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("runs_per_hour",)})
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="window of the 2048^2 workload the CPU baseline times")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="leave out the batched-ensemble summary of the default line")
    ap.add_argument("--no-serving", action="store_true",
                    help="leave out the serving summary of the default line")
    ap.add_argument("--dense-window", type=int, default=300,
                    help="steps over which the default line times the dense ensemble paths")
    ap.add_argument("--precision", choices=sorted(ELEM), default="fp64",
                    help="fp64: the reference's arithmetic (the headline); fp32: FP32-bulk")
    ap.add_argument("--dense", action="store_true",
                    help="ensemble and serve: compute every tile (dry-tile skipping off) on both sides")
    ap.add_argument("--no-dry-skip", action="store_true",
                    help="leave out the second run of the landslide workload with dry-tile skipping")
    ap.add_argument("--no-shard", action="store_true",
                    help="leave out the run of one GPU's strip of the grid cut into 8")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after the timed steps, copy the 7 state planes once device to device "
                         "(a plain copy kernel of known size, for HBM counter calibration)")
    ap.add_argument("--workload", choices=["landslide", "ensemble", "serve", "accuracy"],
                    default="landslide",
                    help="landslide: the 2048^2 FP64 throughput case with the ensemble, serving "
                         "and CPU summaries (default); ensemble, serve: one of those on its own "
                         "line; accuracy: the FP32-against-FP64 table")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=3000)
    ap.add_argument("--floor-members", type=int, default=16)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--max-steps", type=int, default=8000)
    ap.add_argument("--tick", type=int, default=100)
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="GB")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the multi-GPU path is a design, not built); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "ensemble":
        return ensemble(args)
    if args.workload == "serve":
        return serve(args)
    if args.workload == "accuracy":
        return accuracy(args)
    return landslide(args)


def dry_skip_run(sc, args, u_dense, dense_value) -> dict:
    """The same workload with dry-tile skipping (m2p_params.dry_skip = 1); its final state
    must equal the dense run's bit for bit."""
    import numpy as np
    from moses_2pdf_amd import Simulation
    with Simulation(sc, precision=args.precision, dry_skip=True) as sim:
        sim.step(args.warmup + profiled_steps(args))   # the dense run's untimed steps
        sim.sync()
        t0 = time.perf_counter()
        sim.step(args.steps)
        sim.sync()
        el = time.perf_counter() - t0
        u = sim.state()
    v = sc.nx * sc.ny * args.steps / el
    print(f"# dry-skip run: {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    return {"value": v, "unit": "cell-updates/s", "ms_per_step": el / args.steps * 1e3,
            "speedup": v / dense_value, "identical_state": bool(np.array_equal(u, u_dense)),
            "wet_cells": int(((u[0] > 0) | (u[3] > 0)).sum()),
            "note": "same workload and steps with m2p_params.dry_skip = 1: tiles whose inputs "
                    "are all zero are left out (exact); not the headline, which computes every "
                    "tile as the reference does"}


def shard_run(sc, args, dense_value, parts: int = 8) -> dict:
    """One GPU's share of the same grid cut into ``parts`` row strips (nx x ny/parts plus the
    3-cell ghost ring, which stands in for the halo a neighbour would send), every tile
    computed.  The strip is the one holding the most released material.  The halo exchange
    and the dt all-reduce of a multi-GPU run are not in this figure."""
    import dataclasses
    import numpy as np
    from moses_2pdf_amd import Simulation
    h = sc.ny // parts
    rows = sc.depth.sum(axis=1)
    r0 = int(np.argmax(np.convolve(rows, np.ones(h), mode="valid")))
    sub = dataclasses.replace(sc, topo=sc.topo[r0:r0 + h].copy(), depth=sc.depth[r0:r0 + h].copy(),
                              phi=sc.phi[r0:r0 + h].copy())
    with Simulation(sub, precision=args.precision, dry_skip=False) as sim:
        sim.step(args.warmup)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(args.steps)
        sim.sync()
        el = time.perf_counter() - t0
    v = sub.nx * sub.ny * args.steps / el
    print(f"# shard run {sub.nx}x{sub.ny}: {args.steps} steps in {el:.3f} s", file=sys.stderr,
          flush=True)
    return {"grid": [sub.nx, sub.ny], "rows": [r0, r0 + h], "value": v, "unit": "cell-updates/s",
            "ms_per_step": el / args.steps * 1e3, "per_cell_rate_vs_full_grid": v / dense_value,
            "note": f"one GPU's share of the {sc.nx}x{sc.ny} grid cut into {parts} row strips, every "
                    "tile computed; halo exchange and dt all-reduce not included"}


def profiled_steps(args) -> int:
    """Untimed steps after the warmup with HIP events around every kernel family."""
    return min(args.steps, 20)


def limiter(kernel: str, args) -> dict:
    """What holds `kernel` below its HBM roofline, from the committed PMC counters."""
    v = valu_from_profiles(kernel, args.n, args.precision)
    why = ("every FP64 divide and square root stays the IEEE sequence (bit parity with the "
           "oracle): about half of the VALU time, with the quarter-rate reciprocal; the vector "
           "ALU and its dependent FP64 chains at 4 waves per SIMD, not bytes, set the time"
           if args.precision == "fp64" else
           "FP32 arithmetic with hardware reciprocals; VALU issue sets the time")
    return {"kind": "valu", "valu_issue_share": v[0] if v else None,
            "valu_busy_share": v[2] if v else None, "wait_share": v[1] if v else None,
            "source": f"profiles/valu_counters_{args.n}.txt" if v else None, "why": why}


def landslide(args) -> int:
    from moses_2pdf_amd import Simulation
    from moses_2pdf_amd import scenarios as S
    sc = S.landslide_fractal(n=args.n, seed=2)
    print(f"# workload ready: {sc.nx}x{sc.ny}", file=sys.stderr, flush=True)
    # the headline computes every tile, as the reference does; the dry-skipping rate of the
    # same workload is reported beside it
    sim = Simulation(sc, precision=args.precision, dry_skip=False)
    alg_bytes = {k: v * ELEM[args.precision] for k, v in PLANES.items()}
    sim.step(args.warmup)
    sim.sync()
    print(f"# warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    # every kernel family timed with HIP events over an untimed window (the breakdown and
    # the dominant kernel); the timed window then carries events around the dominant
    # kernel only (each event pair between two kernels costs a few microseconds)
    families = ("inflow", "ghost", "cfl", "dt", "stage0", "stage1", "vmax")
    sim.profile(True)
    sim.step(profiled_steps(args))
    kern = {k: sim.profile_get(k) for k in ("ghost", "cfl", "dt", "stage0", "stage1")}
    totals = {k: v[0] * v[1] for k, v in kern.items()}
    dom = max(("cfl", "stage0", "stage1"), key=lambda k: totals[k])
    sim.profile(2 + families.index(dom))
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    dom_ms, dom_launches = sim.profile_get(dom)
    sim.profile(False)
    assert dom_launches == args.steps
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

    avg_ms = dom_ms
    alg = alg_bytes[dom] * cells
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(dom, args.n, args.precision)
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
                     "alg_bytes_per_cell": alg_bytes[dom], "launch_ms": avg_ms,
                     "measured_GBs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                     "limiter": limiter(dom, args),
                     "note": "bound names the roofline the fraction is priced against (HBM: the "
                             "path is a stencil, no MFMA); limiter names what the counters show "
                             "holds the kernel below it"},
        "step_hbm_alg_GBs": step_alg * args.steps / elapsed / 1e9,
        "step_alg_frac": step_alg * args.steps / elapsed / 1e9 / HBM_PEAK_GBS,
        "kernels_ms": {k: v[0] for k, v in kern.items()},
        "kernels_ms_note": "each family timed with HIP events over 20 untimed steps; the "
                           "dominant kernel's figure in roofline comes from the timed steps",
        "state_finite": ok,
        "mass": [ms, mf],
    }
    if not args.no_dry_skip:
        out["dry_skip"] = dry_skip_run(sc, args, u, value)
    if not args.no_shard:
        out["shard"] = shard_run(sc, args, value)
    del sim
    if args.precision == "fp64" and args.n == 2048 and not args.no_ensemble:
        print("# ensemble ...", file=sys.stderr, flush=True)
        r = ensemble_measure(args, dense_steps=args.dense_window)
        out["ensemble"] = dict(workload=ENSEMBLE_WORKLOAD.format(n=r["members"], steps=r["steps"]),
                               **r)
    if args.precision == "fp64" and args.n == 2048 and not args.no_serving:
        print("# serving ...", file=sys.stderr, flush=True)
        r = serve_measure(args)
        out["serving"] = dict(workload=SERVE_WORKLOAD.format(k=r["requests"], rate=args.rate,
                                                             ms=args.max_steps, gb=args.hbm_budget),
                              **r)
    if not args.no_cpu_baseline:
        print("# cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
