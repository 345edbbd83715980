"""Batched-ensemble runs: the plain one-at-a-time path, batched runs in either precision,
and the FP32-versus-FP64 final-depth error at equal simulated time.

Every run returns its wall time and, per member, the final state and simulated time, so
the caller can price scenarios/hour and compare precisions, and the member's non-finite
flag (m2p_nonfinite), so that runs which blew up are reported as failed rather than
compared.
"""
from __future__ import annotations

import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import scenarios as S
from .accuracy import depth_error, flow_depth
from .engine import Batch, Simulation


def log(*a):
    print("#", *a, file=sys.stderr, flush=True)


def open_all(scs, precision, dry_skip=True, workers=None):
    """Opens every scenario, on up to `workers` host threads (default min(8, CPUs)).
    Distinct contexts may be opened from different threads (the C ABI's contract), and
    ctypes releases the GIL inside m2p_open_grid, so the members' host-side padding,
    uploads and set-up kernels overlap.  Returns the simulations in the order of scs; if
    one fails, the others are closed and the error raised."""
    n = workers or min(8, os.cpu_count() or 1)
    if n <= 1 or len(scs) <= 1:
        return [Simulation(sc, precision=precision, dry_skip=dry_skip) for sc in scs]

    def one(sc):
        try:
            return Simulation(sc, precision=precision, dry_skip=dry_skip)
        except Exception as e:   # returned, so that the others still get closed
            return e

    with ThreadPoolExecutor(max_workers=n) as pool:
        sims = list(pool.map(one, scs))
    bad = [x for x in sims if isinstance(x, Exception)]
    if bad:
        for x in sims:
            if not isinstance(x, Exception):
                x.close()
        raise bad[0]
    return sims


def population(count: int = 256, seed: int = 4):
    ms = S.ensemble_members(count, seed)
    return ms, [S.ensemble_scenario(m) for m in ms]


def run_plain(scs, steps, precision="fp64", keep=True, dry_skip=True):
    """Each scenario alone, one after another, `steps` m2p_step steps each.  Returns the
    wall time ("wall_s", of which "open_s" opening the scenarios and "step_s" stepping
    them), each member's (state, time) ("out", if keep) and non-finite flag ("nonfinite")."""
    t0 = time.perf_counter()
    out, bad = [], []
    t_open = t_step = 0.0
    for k, sc in enumerate(scs):
        ta = time.perf_counter()
        with Simulation(sc, precision=precision, dry_skip=dry_skip) as s:
            s.sync()
            tb = time.perf_counter()
            s.step(steps)
            s.sync()
            t_step += time.perf_counter() - tb
            t_open += tb - ta
            if keep:
                out.append((s.state(), s.time_steps()[0]))
            else:
                s.time_steps()
            bad.append(s.nonfinite()[0])
        if (k + 1) % 32 == 0:
            log(f"plain {precision}: {k + 1}/{len(scs)} scenarios, {time.perf_counter() - t0:.1f} s")
    return {"wall_s": time.perf_counter() - t0, "open_s": t_open, "step_s": t_step, "out": out,
            "nonfinite": bad}


def run_batch(scs, steps, precision, tune_steps=3, t_end=None, keep=True, graph_steps=None,
              dry_skip=True):
    """All scenarios in one batch.  With t_end (per member), each member halts at its own
    end time and the batch steps until every member has halted (steps is then a cap).
    "wall_s" runs from the first open to the end of stepping ("open_s", "tune_s" and
    "step_s" are its parts; the rest is adding the members to the batch)."""
    t0 = time.perf_counter()
    sims = open_all(scs, precision, dry_skip)
    t_open = time.perf_counter() - t0
    b = Batch(precision)
    for s in sims:
        b.add(s)
    if t_end is not None:
        for s, te in zip(sims, t_end):
            s.set_stop(t_end=te, max_steps=steps)
    report = ""
    t_tune = 0.0
    if tune_steps:
        t1 = time.perf_counter()
        report = b.tune(tune_steps)
        t_tune = time.perf_counter() - t1
        log("tune:\n" + report)
    if graph_steps is not None:
        b.set("graph_steps", graph_steps)
    b.sync()
    t1 = time.perf_counter()
    if t_end is None:
        left = steps
        while left > 0:
            n = min(500, left)
            b.step(n)
            b.sync()
            left -= n
            log(f"batch {precision}: {steps - left}/{steps} steps, {time.perf_counter() - t1:.1f} s")
    else:
        done = 0
        while done < steps:
            b.step(200)
            b.sync()
            done += 200
            if all(s.status()[2] for s in sims):
                break
            log(f"batch {precision} to t_end: {done} steps")
    t_step = time.perf_counter() - t1
    wall = time.perf_counter() - t0   # opening, adding, tuning and stepping
    out = [(s.state(), s.time_steps()[0]) for s in sims] if keep else []
    nts = [s.status()[1] for s in sims]
    bad = [s.nonfinite()[0] for s in sims]
    b.close()
    for s in sims:
        s.close()
    return {"wall_s": wall, "step_s": t_step, "open_s": t_open, "tune_s": t_tune,
            "tune_report": report, "out": out, "steps": nts, "nonfinite": bad}


def errors(u32s, u64s):
    es = [depth_error(a, b) for a, b in zip(u32s, u64s)]
    l1 = np.array([e["l1"] for e in es])
    li = np.array([e["linf"] for e in es])
    iou = np.array([e["runout_iou"] for e in es])
    return {"members": len(es), "l1_median": float(np.median(l1)), "l1_p90": float(np.percentile(l1, 90)),
            "l1_max": float(l1.max()), "l1_mean": float(l1.mean()),
            "frac_l1_le_1pct": float((l1 <= 0.01).mean()),
            "linf_median": float(np.median(li)), "linf_max": float(li.max()),
            "runout_iou_median": float(np.median(iou)), "runout_iou_min": float(iou.min())}, l1


def floor_l1(scs, steps, ref, count=16, seed=1):
    """FP64 sensitivity floor on a subset: L1 between the FP64 run and an FP64 run from an
    initial state perturbed by 1e-7 relative, at the same time."""
    idx = np.linspace(0, len(scs) - 1, min(count, len(scs))).astype(int)
    out = []
    for k in idx:
        with Simulation(scs[k]) as p:
            u = p.state()
            p.set_state(u * (1 + 1e-7 * np.random.default_rng(seed).standard_normal(u.shape)))
            p.run_to(ref[k][1], 20 * steps)
            out.append(depth_error(p.state(), ref[k][0])["l1"])
    return idx, np.array(out)
