"""Host-side logic of the serving loop and the ensemble runs (CPU, no GPU calls)."""
import math

import numpy as np
import pytest

from moses_2pdf_amd import scenarios as S
from moses_2pdf_amd import serving as V
from moses_2pdf_amd.ensemble import errors


def test_request_stream_statistics():
    # the distributions make_requests draws from (the full scenarios are slow to build)
    rng = np.random.default_rng(11)
    gaps, sides, modes = [], [], []
    for _ in range(4000):
        gaps.append(rng.exponential(1 / 20.0))
        sides.append(int(np.clip(round(384 * math.exp(0.5 * rng.standard_normal())), 128, 2048)))
        modes.append(1 if rng.uniform() < 0.3 else 0)
    assert abs(np.mean(gaps) - 0.05) < 0.005
    assert abs(np.median(sides) - 384) < 20 and min(sides) >= 128 and max(sides) <= 2048
    assert abs(np.mean(modes) - 0.3) < 0.03


def test_make_requests_small():
    reqs = V.make_requests(6, rate=100.0, seed=3, median=64, sigma=0.3, lo=48, hi=96, p_inflow=0.5)
    assert [r.rid for r in reqs] == list(range(6))
    assert all(b.arrival > a.arrival for a, b in zip(reqs, reqs[1:]))
    for r in reqs:
        assert 48 <= r.n <= 96 and r.scenario.nx == r.n and r.scenario.ny == r.n
        assert r.scenario.mode == r.mode
        if r.mode == 1:
            assert r.scenario.inflow_h.shape[1] == r.scenario.inflow_loc.shape[0]
        else:
            assert r.scenario.depth.max() > 0


def test_summary_percentiles():
    res = [V.Result(k, arrival=0.0, start=0.0, done=float(k + 1), steps=100, time=1.0)
           for k in range(100)]
    s = V.summary(res, wall=50.0, max_steps=8000)
    assert s["requests"] == 100 and s["runs_per_hour"] == 100 / 50.0 * 3600
    assert s["latency_p50_s"] == np.percentile(np.arange(1, 101), 50)
    assert s["latency_p99_s"] == np.percentile(np.arange(1, 101), 99)
    assert s["stopped_by_speed"] == 100


def test_ensemble_error_aggregate():
    def st(h):
        u = np.zeros((7, 10, 10))
        u[0, 3:-3, 3:-3] = h
        return u
    ref = [st(np.full((4, 4), 1.0)), st(np.full((4, 4), 2.0))]
    got = [st(np.full((4, 4), 1.005)), st(np.full((4, 4), 2.0))]
    agg, l1 = errors(got, ref)
    assert agg["members"] == 2 and abs(l1[0] - 0.005) < 1e-12 and l1[1] == 0
    assert agg["frac_l1_le_1pct"] == 1.0 and agg["l1_max"] == l1[0]


def test_ensemble_scenario_release_inside_domain():
    for m in S.ensemble_members(16, seed=9):
        if m.n > 256:
            continue
        sc = S.ensemble_scenario(m)
        d = sc.depth
        assert d.max() > 0
        assert d[0].max() == 0 and d[-1].max() == 0 and d[:, 0].max() == 0 and d[:, -1].max() == 0
        vol = d.sum() * sc.cellsize ** 2
        assert abs(vol - m.volume) < 0.1 * m.volume


def test_open_all_keeps_order_and_closes_the_others_when_one_fails(monkeypatch):
    # ensemble.open_all opens on host threads: results come back in the scenarios' order,
    # and when one open raises, every member opened is closed and that error raised
    from moses_2pdf_amd import ensemble as EN

    closed = []

    class FakeSim:
        def __init__(self, sc, precision="fp64", dry_skip=True):
            if sc == "bad":
                raise RuntimeError("open failed")
            self.sc = sc

        def close(self):
            closed.append(self.sc)

    monkeypatch.setattr(EN, "Simulation", FakeSim)
    sims = EN.open_all(list(range(40)), "fp32", workers=8)
    assert [s.sc for s in sims] == list(range(40))
    with pytest.raises(RuntimeError, match="open failed"):
        EN.open_all([0, 1, "bad", 3, 4], "fp32", workers=4)
    assert sorted(closed) == [0, 1, 3, 4]
