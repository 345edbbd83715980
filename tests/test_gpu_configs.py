"""The batched-ensemble and serving configurations at their full sizes (GPU).

* Ensemble: the 256-member population (sides 128/256/512/1024 with weights 4:3:2:1,
  seeded fractal DEMs, Latin-hypercube parameters), autotuned and graph-captured, for all
  3000 steps, in FP64 and in FP32.  Every member's non-finite flag must say exactly
  whether its state is finite; sampled members (two of each size) must be bit-identical
  to the same scenario stepped alone; and a 128^2 member of the FP64 batch must equal the
  CPU oracle after all 3000 steps, bit for bit.
* Serving: the 64-request Poisson stream (20 req/s, sides lognormal around 384 clipped to
  128-2048, landslide/inflow 70/30, stop at a flow speed below 0.01 m/s after motion or
  8000 steps), continuous batching in FP32 in real time; every request is answered, and
  sampled requests (each size class, both modes) give bit-identical results alone.
"""
import sys

import numpy as np
import pytest

from moses_2pdf_amd import Batch, Simulation
from moses_2pdf_amd import ensemble as E
from moses_2pdf_amd import serving as V
from oracle.oracle import Oracle
from util import compare, interior

pytestmark = pytest.mark.gpu

STEPS = 3000


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def population():
    ms, scs = E.population(256, seed=4)
    return ms, scs


def _sample(ms, per_size=2):
    out = []
    for n in (128, 256, 512, 1024):
        idx = [k for k, m in enumerate(ms) if m.n == n]
        out += [idx[0], idx[len(idx) // 2]][:per_size]
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ensemble_256_members_3000_steps(population, precision):
    ms, scs = population
    sims = [Simulation(sc, precision=precision) for sc in scs]
    try:
        with Batch(precision) as b:
            for s in sims:
                b.add(s)
            say(b.tune(3).strip().split("\n")[-1])
            assert b.get("graph_steps") > 0
            for k in range(STEPS // 500):
                b.step(500)
                b.sync()
                say(f"{precision} batch: {(k + 1) * 500} steps")
            states = [s.state() for s in sims]
            times = [s.time_steps() for s in sims]
            b.step(1)      # the next CFL pass reads each member's state after 3000 steps
            b.sync()
            flags = [s.nonfinite() for s in sims]
        bad = [k for k in range(len(scs)) if not np.isfinite(states[k]).all()]
        say(f"{precision}: non-finite members {bad}, flags {[k for k, f in enumerate(flags) if f[0]]}")
        for k in range(len(scs)):
            assert flags[k][0] == (k in bad), k
            if flags[k][0]:
                assert 0 < flags[k][1] <= STEPS
            assert times[k][1] == STEPS
        assert len(bad) <= 8
        for k in _sample(ms):
            with Simulation(scs[k], precision=precision) as a:
                a.step(STEPS)
                assert a.time_steps() == times[k], k
                assert np.array_equal(a.state().view(np.uint64), states[k].view(np.uint64)), k
        if precision == "fp64":
            import os
            from oracle import oracle as O
            O.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
            k = next(k for k, m in enumerate(ms) if m.n == 128 and k not in bad)
            o = Oracle(scs[k])
            o.step(STEPS)
            rel, ndiff = compare(interior(states[k]), interior(o.get("u")))
            say(f"member {k} (128^2) after {STEPS} steps: batch vs oracle differing {ndiff}")
            assert ndiff == 0 and times[k][0] == o.scalar("time")
    finally:
        for s in sims:
            s.close()


def test_serving_64_requests_mix_equals_alone():
    reqs = V.make_requests(64, rate=20.0, seed=5)
    assert min(r.n for r in reqs) >= 128 and max(r.n for r in reqs) <= 2048
    srv = V.Server(precision="fp32", tick_steps=100, max_steps=8000, vstop=0.01, keep_state=True)
    res = srv.run(reqs, realtime=True)
    wall = max(r.done for r in res)
    summ = V.summary(res, wall, 8000)
    say("serving:", {k: summ[k] for k in ("runs_per_hour", "latency_p50_s", "latency_p99_s",
                                           "stopped_by_speed", "failed_nonfinite")},
        "peak active", srv.peak_active)
    assert [r.rid for r in res] == list(range(64))
    assert srv.peak_active >= 8                  # requests overlapped on the device
    for r in res:
        assert r.steps <= 8000
        assert r.failed or np.isfinite(r.state).all()
    order = sorted(range(64), key=lambda k: reqs[k].n)
    pick = sorted({order[0], order[16], order[32], order[48], order[63],
                   next(k for k in range(64) if reqs[k].mode == 1),
                   next(k for k in range(64) if reqs[k].mode == 0)})
    solo = V.serve_alone([reqs[k] for k in pick], precision="fp32", max_steps=8000,
                         tick_steps=100, vstop=0.01, keep_state=True, realtime=False)
    for k, b in zip(pick, solo):
        a = res[k]
        say(f"request {k}: side {reqs[k].n} mode {reqs[k].mode} steps {a.steps}")
        assert (a.steps, a.time, a.failed) == (b.steps, b.time, b.failed), k
        if not a.failed:
            assert np.array_equal(a.state.view(np.uint64), b.state.view(np.uint64)), k
