"""Batched ensembles: every member of a batch ends bit-identical to the same scenario
stepped alone (whose FP64 path is bit-identical to the oracle), for both precisions,
every compiled stage-kernel variant and block schedule, graph and non-graph stepping,
members of different sizes and modes, members halting on their own rules, and members
leaving mid-run.  The autotuner restores every member's state."""
import sys

import numpy as np
import pytest

from moses_2pdf_amd import Batch, Simulation
from moses_2pdf_amd import scenarios as S
from util import df004

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def population():
    ms = S.ensemble_members(256)
    mem = [S.ensemble_scenario(m) for m in [m for m in ms if m.n == 128][:3]]
    mem.append(S.landslide_plane(n=72))
    mem.append(S.landslide_fractal(n=40, seed=9))          # per-cell phi
    sc = S.landslide_plane(n=50)
    sc.cfl, sc.rho_ratio = 0.4, 1000.0 / 2700.0
    mem.append(sc)
    mem.append(df004())                                      # inflow mode
    return mem


def alone(scs, steps, precision, metric_form=0):
    out = []
    for sc in scs:
        with Simulation(sc, precision=precision, metric_form=metric_form) as s:
            s.step(steps)
            out.append((s.state(), s.time_steps()))
    return out


def batched(scs, steps, precision, chunks=None, setup=None, metric_form=0):
    sims = [Simulation(sc, precision=precision, metric_form=metric_form) for sc in scs]
    b = Batch(precision)
    for s in sims:
        b.add(s)
    if setup:
        setup(b)
    for n in (chunks or [steps]):
        b.step(n)
    b.sync()
    res = [(s.state(), s.time_steps()) for s in sims]
    b.close()
    for s in sims:
        s.close()
    return res


def same(a, b):
    return all(np.array_equal(x[0], y[0]) and x[1] == y[1] for x, y in zip(a, b))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_alone(precision):
    scs = population()
    ref = alone(scs, 37, precision)
    got = batched(scs, 37, precision, chunks=[1, 20, 16])   # graph of 16 steps + direct
    for k, (x, y) in enumerate(zip(got, ref)):
        assert x[1] == y[1], (k, x[1], y[1])
        assert np.array_equal(x[0], y[0]), k
    say(precision, "batch of", len(scs), "equal to alone after 37 steps")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_with_stored_metric_equals_alone(precision):
    """Members that load the stored metric planes (m2p_params.metric_form = 1) make the
    batch take its stored-metric kernels; every member still ends equal to itself stepped
    alone, and in FP64 equal to the formed-metric run as well."""
    scs = population()
    ref = alone(scs, 37, precision, metric_form=1)
    got = batched(scs, 37, precision, chunks=[1, 20, 16], metric_form=1)
    for k, (x, y) in enumerate(zip(got, ref)):
        assert x[1] == y[1], (k, x[1], y[1])
        assert np.array_equal(x[0], y[0]), k
    if precision == "fp64":
        assert same(got, alone(scs, 37, precision))
    say(precision, "stored-metric batch of", len(scs), "equal to alone after 37 steps")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_variant_and_schedule_gives_the_same_result(precision):
    scs = population()[:5]
    ref = alone(scs, 12, precision)
    b0 = Batch(precision)
    from moses_2pdf_amd import engine
    import ctypes
    variants = {"fp64": [(7, 4), (3, 4), (7, 3)], "fp32": [(7, 6), (7, 4), (3, 6), (15, 4), (3, 8), (7, 8)]}[precision]
    for ty, w in variants:
        for remap in (0, 1):
            for order in (0, 1):
                def setup(b, ty=ty, w=w, remap=remap, order=order):
                    b.set("ty", ty)
                    b.set("waves", w)
                    b.set("remap", remap)
                    b.set("order", order)
                    b.set("graph_steps", 5)
                got = batched(scs, 12, precision, setup=setup)
                assert same(got, ref), (ty, w, remap, order)
    b0.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cfl_block_count_gives_the_same_result(precision):
    # each box keeps its own partial slot, so the number of blocks striding over the boxes
    # (from 64, many boxes per block, to more blocks than boxes) cannot change a result
    scs = population()
    ref = alone(scs, 15, precision)
    for cb in (64, 1000, 32768):
        def setup(b, cb=cb):
            b.set("cfl_blocks", cb)
            assert b.get("cfl_blocks") == cb
        got = batched(scs, 15, precision, chunks=[3, 12], setup=setup)
        assert same(got, ref), cb
    say(precision, "cfl_blocks 64 / 1000 / 32768 equal to alone")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plan_cap_gives_the_same_result(precision):
    # with dry-tile skipping the stage launch takes the first plan_cap entries of each tile
    # list and a second launch the rest: at plan_cap 1 nearly every tile goes through the
    # second (its loop over entries), at 1024 none does
    scs = population()
    ref = alone(scs, 15, precision)
    for cap in (1, 5, 1024):
        def setup(b, cap=cap):
            b.set("plan_cap", cap)
            b.set("graph_steps", 4)
            assert b.get("plan_cap") == cap
        got = batched(scs, 15, precision, chunks=[3, 12], setup=setup)
        assert same(got, ref), cap
    say(precision, "plan_cap 1 / 5 / 1024 equal to alone")


def test_tune_restores_state_and_keeps_results():
    scs = population()
    ref = alone(scs, 10, "fp32")
    sims = [Simulation(sc, precision="fp32") for sc in scs]
    b = Batch("fp32")
    for s in sims:
        b.add(s)
    b.step(4)
    b.sync()
    before = [s.state() for s in sims]
    rep = b.tune(trial_steps=2)
    say(rep)
    assert "chosen:" in rep
    after = [s.state() for s in sims]
    assert all(np.array_equal(x, y) for x, y in zip(before, after))
    assert all(s.time_steps()[1] == 4 for s in sims)
    b.step(6)
    b.sync()
    assert all(np.array_equal(s.state(), r[0]) for s, r in zip(sims, ref))
    b.close()


def test_tune_with_halted_and_inflow_members():
    # the tuner steps the batch once before its trials and puts every member back after:
    # a member halted before the tune stays halted at its step, an inflow member keeps its
    # hydrograph position, and the others continue as if the tune had not run
    scs = population()
    sims = [Simulation(sc) for sc in scs]
    b = Batch("fp64")
    for s in sims:
        b.add(s)
    sims[0].set_stop(max_steps=3)
    b.step(5)
    b.sync()
    before = [(s.state(), s.status()) for s in sims]
    assert before[0][1][1] == 3 and before[0][1][2]
    b.tune(trial_steps=1)
    after = [(s.state(), s.status()) for s in sims]
    for (u0, st0), (u1, st1) in zip(before, after):
        assert st0 == st1
        assert np.array_equal(u0.view(np.uint64), u1.view(np.uint64))
    b.step(7)
    b.sync()
    got = [(s.state(), s.time_steps()) for s in sims]
    b.close()
    for s in sims:
        s.close()
    ref = alone(scs[1:], 12, "fp64")
    for k, (x, y) in enumerate(zip(got[1:], ref)):
        assert x[1] == y[1], k
        assert np.array_equal(x[0], y[0]), k
    with Simulation(scs[0]) as s:
        s.step(3)
        assert np.array_equal(got[0][0], s.state())


def test_members_halt_on_their_own_rules():
    scs = population()[:4]
    sims = [Simulation(sc) for sc in scs]
    ref_t = None
    with Simulation(scs[1]) as s:
        s.step(6)
        t6 = s.time_steps()[0]
        s.step(1)
        ref_t = 0.5 * (t6 + s.time_steps()[0])   # inside step 7
    b = Batch("fp64")
    for s in sims:
        b.add(s)
    sims[0].set_stop(max_steps=5)
    sims[1].set_stop(t_end=ref_t)
    b.step(20)
    b.sync()
    st = [s.status() for s in sims]
    # the batch's one-copy status agrees with each member's own
    for s, row in zip(sims, b.status()):
        assert row[:3] == s.status() and row[3:] == s.nonfinite()
    b.remove(sims[2])
    assert [r[1] for r in b.status()] == [sims[k].status()[1] for k in (0, 1, 3)]
    b.add(sims[2])
    assert st[0][1] == 5 and st[0][2]
    assert st[1][1] == 7 and st[1][0] == ref_t and st[1][2]
    assert st[2][1] == 20 and not st[2][2]
    # the halted members equal their alone runs at the halting point
    with Simulation(scs[0]) as s:
        s.step(5)
        assert np.array_equal(s.state(), sims[0].state())
    with Simulation(scs[1]) as s:
        s.run_to(ref_t)
        assert np.array_equal(s.state(), sims[1].state())
    b.close()


def test_member_leaves_mid_run_and_steps_alone():
    scs = population()[:4]
    ref = alone(scs, 15, "fp64")
    sims = [Simulation(sc) for sc in scs]
    b = Batch("fp64")
    for s in sims:
        b.add(s)
    b.step(5)
    b.remove(sims[2])
    assert len(b) == 3
    with pytest.raises(Exception):
        b.add(sims[0])              # already a member
    sims[2].step(10)                # alone again
    b.step(10)
    b.sync()
    for s, r in zip(sims, ref):
        assert np.array_equal(s.state(), r[0])
    b.close()


def test_stop_speed_rule_halts_a_settling_flow():
    sc = S.landslide_plane(n=48)
    with Simulation(sc) as s:
        s.set_stop(vstop=1e9)          # never reached: never halts
        s.step(30)
        assert not s.status()[2]
        s.set_stop(vstop=1e-30)        # reached at once, then never below it
        s.step(30)
        assert not s.status()[2]


def test_members_opened_on_threads_match_opened_one_by_one():
    # ensemble.open_all opens on several host threads (pooled streams are shared between
    # contexts): each member's set-up must not see another's, so the batch ends with the
    # same bits as members opened one after another
    from moses_2pdf_amd import ensemble as EN
    scs = population() * 3
    ref = batched(scs, 20, "fp32")
    sims = EN.open_all(scs, "fp32", workers=8)
    b = Batch("fp32")
    for s in sims:
        b.add(s)
    b.step(20)
    b.sync()
    got = [(s.state(), s.time_steps()) for s in sims]
    b.close()
    for s in sims:
        s.close()
    for (x, tx), (y, ty) in zip(got, ref):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)) and tx == ty
