"""The serving loop: each request's result is the same alone and in the mix (bit-identical
final state, step count and simulated time), and admission keeps the device footprint of
the scenarios held within the budget."""
import numpy as np
import pytest

from moses_2pdf_amd.engine import footprint
from moses_2pdf_amd.serving import Server, make_requests, serve_alone

pytestmark = pytest.mark.gpu


def small_requests(k=6):
    return make_requests(k, rate=200.0, seed=7, median=96, sigma=0.4, lo=64, hi=160, p_inflow=0.4)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_results_same_alone_and_in_the_mix(precision):
    reqs = small_requests()
    assert {r.mode for r in reqs} == {0, 1}
    srv = Server(precision=precision, tick_steps=20, max_steps=300, vstop=0.01, keep_state=True,
                 graph_steps=5)
    mix = srv.run(reqs, realtime=False)
    solo = serve_alone(reqs, precision=precision, max_steps=300, tick_steps=20, vstop=0.01,
                       keep_state=True, realtime=False)
    assert srv.peak_active >= 2
    for a, b in zip(mix, solo):
        assert a.rid == b.rid
        assert (a.steps, a.time) == (b.steps, b.time), (a.rid, a.steps, b.steps)
        assert np.array_equal(a.state, b.state), a.rid


def test_fp64_mix_results_equal_the_oracle():
    """In FP64 a request served in the mix ends bit-identical to the CPU oracle run of the
    same scenario for the same number of steps (landslide and inflow requests)."""
    import os
    from oracle import oracle as O
    from oracle.oracle import Oracle
    O.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    reqs = small_requests()
    srv = Server(precision="fp64", tick_steps=20, max_steps=200, vstop=0.01, keep_state=True,
                 graph_steps=5)
    mix = srv.run(reqs, realtime=False)
    for mode in (0, 1):
        k = next(i for i, r in enumerate(reqs) if r.mode == mode)
        o = Oracle(reqs[k].scenario)
        o.step(mix[k].steps)
        uo = o.get("u")[:, 3:-3, 3:-3]
        assert np.array_equal(mix[k].state[:, 3:-3, 3:-3], uo), (k, mode)
        assert mix[k].time == o.scalar("time")


def test_admission_keeps_footprint_within_budget():
    reqs = small_requests(8)
    peaks = [footprint(r.scenario.mode, r.n, r.n, "fp32",
                       *(r.scenario.inflow_h.shape if r.mode == 1 else (0, 0)))[1] for r in reqs]
    budget = 2.2 * max(peaks)
    srv = Server(precision="fp32", hbm_budget=budget, tick_steps=20, max_steps=120)
    res = srv.run(reqs, realtime=False)
    assert len(res) == len(reqs)
    assert srv.peak_held <= budget
    assert all(r.done >= r.start >= r.arrival - 1e-9 for r in res)


def test_footprint_grows_with_grid_and_precision():
    a64, p64 = footprint(0, 256, 256, "fp64")
    a32, p32 = footprint(0, 256, 256, "fp32")
    b64, _ = footprint(0, 512, 512, "fp64")
    assert p64 >= a64 and p32 >= a32
    assert b64 > 3.5 * a64
    P = 262 * 262
    assert a64 >= (38 + 16) * 8 * P   # metric (cell and face planes) + state planes


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_footprint_matches_device_memory(prec):
    import ctypes
    from moses_2pdf_amd import Simulation
    from moses_2pdf_amd import scenarios as S
    hip = ctypes.CDLL("libamdhip64.so.7")   # the HIP runtime the library itself uses

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    with Simulation(S.landslide_plane(n=16)):
        pass                              # the runtime is initialised before the first reading
    free0 = free_bytes()
    with Simulation(S.landslide_plane(n=1024), precision=prec):
        free1 = free_bytes()
    steady, peak = footprint(0, 1024, 1024, prec)
    held = free0 - free1
    print(prec, "held", held, "steady", steady, "peak", peak, flush=True)
    assert peak >= steady
    # the HIP runtime may keep the freed set-up temporaries for the process to reuse, so
    # the device can show anything between the steady and the peak figure as held
    assert 0.95 * steady <= held <= 1.05 * peak + (8 << 20)
