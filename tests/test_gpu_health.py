"""Non-finite detection and halting rules through the C ABI (GPU).

The reference drops a NaN wave speed silently (its speed floor and face maxima use fmax,
its dt reduction fmaxf: MoSES_2PDF_landslide/adNOC_2Pb_kernels.cu:1762-1807, 1988, 1996),
so a run that blows up keeps stepping NaNs.  Ensemble member 187 (128^2, seed 400199) is
a scenario on which the scheme itself fails: the oracle's state first holds a non-finite
value after 1033 steps.  The library must report exactly that step, alone and batched,
with every state before it bit-identical to the oracle's; the serving loop must return
such a request as failed."""
import sys

import numpy as np
import pytest

from moses_2pdf_amd import Batch, Simulation
from moses_2pdf_amd import scenarios as S
from moses_2pdf_amd import serving as V
from oracle.oracle import Oracle
from util import compare, interior

pytestmark = pytest.mark.gpu

BLOWUP_MEMBER, BLOWUP_NT = 187, 1033


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _threads():
    import os
    from oracle import oracle as O
    return O.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))


def blowup_scenario():
    return S.ensemble_scenario(S.ensemble_members(256, 4)[BLOWUP_MEMBER])


def test_nonfinite_flag_rises_at_the_oracles_first_nonfinite_state():
    _threads()
    sc = blowup_scenario()
    o = Oracle(sc)
    o.step(BLOWUP_NT - 1)
    assert np.isfinite(o.get("u")).all()
    with Simulation(sc) as g:
        g.step(BLOWUP_NT - 1)
        rel, ndiff = compare(interior(g.state()), interior(o.get("u")))
        say("member 187 at", BLOWUP_NT - 1, "steps: differing values", ndiff)
        assert ndiff == 0
        assert g.nonfinite() == (False, -1)
        o.step(1)
        g.step(1)
        uo, ug = interior(o.get("u")), interior(g.state())
        assert not np.isfinite(uo).all()            # the scheme itself fails here
        assert np.array_equal(np.isfinite(ug), np.isfinite(uo))
        # Away from the cells that went non-finite the states are still identical.  Next to
        # them they may differ: once an input is NaN, the reference's minmod (fmin of
        # products, L/K:1043-1119, which drops a NaN operand and keeps the other) and the
        # exact minmod here (0 unless both differences share a sign) disagree.
        bad = ~np.isfinite(uo).all(axis=0)
        near = bad.copy()
        for _ in range(4):
            near = near | np.roll(near, 1, 0) | np.roll(near, -1, 0)
            near = near | np.roll(near, 1, 1) | np.roll(near, -1, 1)
        assert near.sum() < 0.05 * near.size
        assert np.array_equal(ug[:, ~near], uo[:, ~near])
        assert g.nonfinite() == (False, -1)         # its CFL pass read the finite state
        g.step(1)
        assert g.nonfinite() == (True, BLOWUP_NT)   # the next step's CFL pass sees it
        g.step(5)
        assert g.nonfinite() == (True, BLOWUP_NT)   # it stays, with the first step
        g.set_state(np.nan_to_num(o.get("u"), nan=0.0, posinf=0.0, neginf=0.0))
        assert g.nonfinite() == (False, -1)         # a new state clears it


def test_nonfinite_flag_in_a_batch():
    others = [S.ensemble_scenario(m) for m in S.ensemble_members(256, 4)[:3]]
    sims = [Simulation(sc) for sc in [blowup_scenario()] + others]
    try:
        with Batch("fp64") as b:
            for s in sims:
                b.add(s)
            b.step(BLOWUP_NT + 1)
            b.sync()
            flags = [s.nonfinite() for s in sims]
        say("batch flags", flags)
        assert flags[0] == (True, BLOWUP_NT)
        assert all(f == (False, -1) for f in flags[1:])
        assert all(np.isfinite(s.state()).all() for s in sims[1:])
    finally:
        for s in sims:
            s.close()


def test_serving_returns_a_blown_up_request_as_failed():
    good = S.ensemble_scenario(S.ensemble_members(256, 4)[0])
    reqs = [V.Request(0, 0.0, blowup_scenario(), 0, 128), V.Request(1, 0.0, good, 0, good.nx)]
    srv = V.Server(precision="fp64", tick_steps=100, max_steps=1200, vstop=-1.0, keep_state=True)
    res = srv.run(reqs, realtime=False)
    say([(r.rid, r.steps, r.failed, r.nonfinite_nt) for r in res])
    assert res[0].failed and res[0].nonfinite_nt == BLOWUP_NT and res[0].state is None
    assert res[0].steps < 1200                      # retired at the first tick after it
    assert not res[1].failed and res[1].steps == 1200 and np.isfinite(res[1].state).all()
    assert V.summary(res, 1.0, 1200)["failed_nonfinite"] == 1


def test_run_to_keeps_the_callers_stop_rules():
    sc = S.landslide_plane(n=40)
    with Simulation(sc) as g:
        g.set_stop(max_steps=10)
        assert g.run_to(1e9, 100) == 10             # the caller's budget holds inside run_to
        assert g.status()[2]                        # and the scenario stays halted
        g.step(5)
        assert g.time_steps()[1] == 10
    with Simulation(sc) as g:
        g.set_stop(max_steps=50)
        g.step(3)
        t3 = g.time_steps()[0]
        n = g.run_to(2 * t3, 1000)                  # ends on the call's own time...
        assert g.time_steps()[0] == 2 * t3 and not g.status()[2]
        g.step(100)                                 # ...and the budget is still 50 steps
        assert g.time_steps()[1] == 50 and g.status()[2]
        assert n >= 1
    with Simulation(sc) as g:
        g.step(2)
        t2 = g.time_steps()[0]
        g.set_stop(t_end=3 * t2)
        g.run_to(10 * t2, 1000)                     # the caller's nearer end time wins
        assert g.time_steps()[0] == 3 * t2 and g.status()[2]
    with Simulation(sc) as a, Simulation(sc) as b:  # an end time set before run_to is the same
        a.step(2)                                   # end time as run_to's own
        b.step(2)
        t2 = a.time_steps()[0]
        a.set_stop(t_end=3 * t2)
        a.step(100)
        b.run_to(3 * t2, 100)
        assert a.time_steps() == b.time_steps()
        assert np.array_equal(a.state(), b.state())
