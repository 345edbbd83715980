"""Continuous-batching serving loop for hazard-scenario requests.

Requests arrive with their inputs (DEM, release or hydrograph, parameters).  The server
keeps one batch of a fixed precision on the device.  Between ticks of `tick_steps` batched
steps it admits waiting requests in arrival order while the device footprint of the
scenarios it holds, plus the newcomer's set-up peak (m2p_footprint), stays within
`hbm_budget`.  It retires scenarios whose device-side stop rule (a max flow speed below
`vstop` after motion, or `max_steps` steps) has halted them, and, as failed, scenarios
whose state has gone non-finite (m2p_nonfinite; no result state is handed back for them).
A request's latency runs from its arrival to the moment its result (final state) is on
the host.

Each scenario's halting rule is evaluated on the device from its own state, and the
batched kernels compute every member exactly as if it ran alone, so a result does not
depend on which other requests shared the device (tests/test_gpu_serving.py checks this).
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from . import scenarios as S
from .engine import Batch, Simulation, footprint


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float             # seconds after the start of the run
    scenario: S.Scenario
    mode: int
    n: int


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    start: float               # admitted (set-up begins)
    done: float                # result on the host
    steps: int
    time: float                # simulated time reached (scaled)
    state: Optional[np.ndarray] = None
    failed: bool = False       # the state went non-finite (no result)
    nonfinite_nt: int = -1     # step count of the first non-finite state seen

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_scenario(rng: np.random.Generator, n: int, mode: int, seed: int) -> S.Scenario:
    """A landslide on a fractal DEM (ensemble sampling) or a debris-flow inflow down a
    V-shaped valley, on an n x n grid at 20 m."""
    delta0 = rng.uniform(15.0, 30.0)
    phi = rng.uniform(0.4, 0.7)
    if mode == 0:
        vol = math.exp(rng.uniform(math.log(1e5), math.log(5e6)))
        m = S.MemberParams(n, seed, delta0, rng.uniform(30.0, 40.0), vol, phi)
        return S.ensemble_scenario(m)
    sc = S.inflow_valley(nx=n, ny=n, seed=seed, nloc=max(4, n // 64))
    sc.delta0 = delta0
    sc.phiS0 = phi
    return sc


def make_requests(count: int, rate: float = 20.0, seed: int = 5, median: int = 384,
                  sigma: float = 0.5, lo: int = 128, hi: int = 2048,
                  p_inflow: float = 0.3) -> List[Request]:
    """Poisson arrivals at `rate` per second; grid side lognormal (median, sigma of the
    log), clipped to [lo, hi]; landslide or inflow with probability 1 - p_inflow, p_inflow."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(np.clip(round(median * math.exp(sigma * rng.standard_normal())), lo, hi))
        mode = 1 if rng.uniform() < p_inflow else 0
        out.append(Request(k, t, request_scenario(rng, n, mode, seed * 1000003 + k), mode, n))
    return out


def _peak(req: Request, precision: str) -> int:
    sc = req.scenario
    nsamp, nloc = (sc.inflow_h.shape if sc.mode == 1 else (0, 0))
    return footprint(sc.mode, sc.nx, sc.ny, precision, nsamp, nloc)


class Server:
    def __init__(self, precision: str = "fp32", hbm_budget: float = 200e9, tick_steps: int = 100,
                 vstop: float = 0.01, max_steps: int = 8000, max_active: int = 256,
                 keep_state: bool = False, graph_steps: int = 25, dry_skip: bool = True):
        self.precision = precision
        self.dry_skip = dry_skip
        self.hbm_budget = hbm_budget
        self.tick_steps = tick_steps
        self.vstop = vstop
        self.max_steps = max_steps
        self.max_active = max_active
        self.keep_state = keep_state
        self.graph_steps = graph_steps
        self.peak_held = 0      # largest device footprint held at once (bytes)
        self.peak_active = 0

    def run(self, requests: List[Request], realtime: bool = True) -> List[Result]:
        b = Batch(self.precision)
        b.set("graph_steps", self.graph_steps)
        queue = list(requests)
        active = {}             # rid -> (Simulation, Request, start, steady bytes)
        held = 0
        results = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while queue or active:
            # admission, in arrival order, while the footprint fits
            while queue and queue[0].arrival <= now() and len(active) < self.max_active:
                req = queue[0]
                steady, peak = _peak(req, self.precision)
                if held + peak > self.hbm_budget and active:
                    break
                queue.pop(0)
                start = now()
                sim = Simulation(req.scenario, precision=self.precision, dry_skip=self.dry_skip)
                sim.set_stop(vstop=self.vstop, max_steps=self.max_steps)
                b.add(sim)
                held += steady
                active[req.rid] = (sim, req, start, steady)
                self.peak_held = max(self.peak_held, held)
                self.peak_active = max(self.peak_active, len(active))
            if not active:
                if queue and realtime:
                    time.sleep(max(0.0, queue[0].arrival - now()))
                elif queue:
                    t0 -= queue[0].arrival - now()   # jump the clock to the next arrival
                continue
            b.step(self.tick_steps)
            b.sync()
            rid_of = {id(sim): rid for rid, (sim, _, _, _) in active.items()}
            # every member's status with one copy (m2p_batch_status)
            for sim, (t, nt, halted, bad, bad_nt) in zip(list(b.members), b.status()):
                rid = rid_of[id(sim)]
                sim, req, start, steady = active[rid]
                if halted or bad:
                    st = sim.state() if (self.keep_state and not bad) else None
                    results.append(Result(rid, req.arrival, start, now(), nt, t, st, bad, bad_nt))
                    b.remove(sim)
                    sim.close()
                    held -= steady
                    del active[rid]
        b.close()
        return sorted(results, key=lambda r: r.rid)


def serve_alone(requests: List[Request], precision: str = "fp64", vstop: float = 0.01,
                max_steps: int = 8000, tick_steps: int = 100, realtime: bool = True,
                keep_state: bool = False, dry_skip: bool = True) -> List[Result]:
    """The one-at-a-time path: requests in arrival order, each alone on the device with
    m2p_step, the same stop rule checked every tick_steps steps."""
    results = []
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    for req in requests:
        if req.arrival > now():
            if realtime:
                time.sleep(req.arrival - now())
            else:
                t0 -= req.arrival - now()
        start = now()
        with Simulation(req.scenario, precision=precision, dry_skip=dry_skip) as sim:
            sim.set_stop(vstop=vstop, max_steps=max_steps)
            while True:
                sim.step(tick_steps)
                t, nt, halted = sim.status()
                bad, bad_nt = sim.nonfinite()
                if halted or bad:
                    break
            st = sim.state() if (keep_state and not bad) else None
        results.append(Result(req.rid, req.arrival, start, now(), nt, t, st, bad, bad_nt))
    return results


def summary(results: List[Result], wall: float, max_steps: int = 8000) -> dict:
    lat = np.array([r.latency for r in results])
    return {"requests": len(results), "wall_s": wall, "runs_per_hour": len(results) / wall * 3600.0,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()),
            "steps_mean": float(np.mean([r.steps for r in results])),
            "stopped_by_speed": int(sum(1 for r in results if r.steps < max_steps and not r.failed)),
            "failed_nonfinite": int(sum(1 for r in results if r.failed))}
