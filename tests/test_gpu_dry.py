"""Dry-tile skipping (m2p_params.dry_skip).  A run that leaves out the tiles whose inputs
are all zero must end bit-identical -- every state value, the simulated time and the step
count -- to the same run computing every tile (which the parity tests tie to the oracle).
Covered: both precisions, landslide and inflow mode, the reference's Hsiaolin data, a
batch, and the state changes after which the flags are set again from the state
(set_state, a checkpoint restore, switching the skipping off and on mid-run).  On a mostly dry grid the stage
kernels must also get much cheaper, or nothing is being skipped."""
import os
import sys

import numpy as np
import pytest

from moses_2pdf_amd import Batch, Simulation
from moses_2pdf_amd import scenarios as S
from util import df004, hsiaolin

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(sc, steps, precision="fp64", dry=True):
    with Simulation(sc, precision=precision, dry_skip=dry) as s:
        s.step(steps)
        return s.state(), s.time_steps(), s.mass()


def same(a, b):
    # bit patterns, so that a -0 where the other run has +0 counts as a difference
    ua, ta, ma = a
    ub, tb, mb = b
    assert ta == tb
    ba, bb = ua.view(np.uint64), ub.view(np.uint64)
    assert np.array_equal(ba, bb), int((ba != bb).sum())
    assert ma == mb


CASES = {
    "fractal256": (lambda: S.landslide_fractal(n=256, seed=2), 150),
    "plane200": (lambda: S.landslide_plane(n=200), 150),
    "hsiaolin": (hsiaolin, 60),
    "df004": (df004, 80),
}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_dry_skip_is_bit_identical(case, precision):
    make, steps = CASES[case]
    sc = make()
    on = run(sc, steps, precision, True)
    off = run(sc, steps, precision, False)
    say(f"{case} {precision}: {steps} steps, t = {on[1][0]:.6f}, wet cells "
        f"{int((on[0][0] > 0).sum())}")
    same(on, off)


def test_dry_skip_switched_mid_run():
    sc = S.landslide_fractal(n=256, seed=5)
    ref = run(sc, 90, dry=False)
    with Simulation(sc, dry_skip=True) as s:
        s.step(30)
        s.set_dry_skip(False)
        s.step(30)
        s.set_dry_skip(True)
        s.step(30)
        same((s.state(), s.time_steps(), s.mass()), ref)


def test_dry_skip_after_set_state_and_restore(tmp_path):
    sc = S.landslide_fractal(n=256, seed=6)

    def with_set_state(dry):
        with Simulation(sc, dry_skip=dry) as s:
            s.step(40)
            u = s.state()
            # a state with the flow moved: the flags must be rebuilt from it
            s.set_state(np.roll(u, 5, axis=2))
            s.step(40)
            return s.state(), s.time_steps(), s.mass()

    same(with_set_state(True), with_set_state(False))

    ck = os.path.join(str(tmp_path), "dry.ckpt")
    with Simulation(sc, dry_skip=True) as a:
        a.step(40)
        a.checkpoint(ck)
    with Simulation(sc, dry_skip=True) as b:
        b.step(7)            # flags of another state, replaced by the restore
        b.restore(ck)
        b.step(40)
        got = (b.state(), b.time_steps(), b.mass())
    same(got, run(sc, 80, dry=False))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dry_skip_in_a_batch(precision):
    scs = [S.landslide_fractal(n=256, seed=11), S.landslide_plane(n=130), df004()]
    ref = [run(sc, 60, precision, False) for sc in scs]
    sims = [Simulation(sc, precision=precision, dry_skip=True) for sc in scs]
    b = Batch(precision)
    for s in sims:
        b.add(s)
    b.tune(2)             # trial steps on every variant, then the state restored
    b.step(60)
    b.sync()
    got = [(s.state(), s.time_steps(), s.mass()) for s in sims]
    b.close()
    for s in sims:
        s.close()
    for g, r in zip(got, ref):
        same(g, r)


def test_stop_speed_rule_same_with_and_without_skipping():
    # 520^2: the dense stages take 7-row tiles and the skipping ones 3-row tiles, so the
    # flow-speed reduction behind the stop speed must read the partials of the tiling that
    # ran; the flow (rows 248-278) lies in tiles that only the 3-row numbering reaches past
    # the 7-row tile count.  At this stop speed the flow reaches it and falls below it within
    # a few steps.
    sc = S.landslide_plane(n=520)
    got = []
    for dry in (False, True):
        with Simulation(sc, dry_skip=dry) as s:
            s.set_stop(vstop=2.0, max_steps=3000)
            s.step(3000)
            got.append((s.status(), s.state()))
    (sa, ua), (sb, ub) = got
    say(f"stop speed: dense {sa}, skipping {sb}")
    assert sa == sb and sa[2] and sa[1] < 3000
    assert np.array_equal(ua.view(np.uint64), ub.view(np.uint64))


def test_dry_skip_skips():
    sc = S.landslide_fractal(n=1024, seed=2)   # 0.05% of the cells start wet
    ms = {}
    for dry in (False, True):
        with Simulation(sc, dry_skip=dry) as s:
            s.step(3)
            s.sync()
            s.profile(True)
            s.step(20)
            s.sync()
            ms[dry] = sum(s.profile_get(k)[0] for k in ("cfl", "stage0", "stage1"))
    say(f"cfl + stages per step: {ms[False]:.3f} ms computing every tile, "
        f"{ms[True]:.3f} ms skipping dry tiles")
    assert ms[True] < 0.4 * ms[False]
