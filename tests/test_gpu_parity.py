"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same inputs.

The bar: <= 1e-10 relative on depths and momenta, total solid and fluid mass equal to
1e-12 relative.  The implementation keeps the reference's operation order
with FMA contraction off, so the tests also report how many values differ at all; the
expected count is 0.
"""
import math
import sys

import numpy as np
import pytest

from moses_2pdf_amd import Simulation
from moses_2pdf_amd import scenarios as S
from oracle.oracle import Oracle
from util import compare, df004, hsiaolin, interior

pytestmark = pytest.mark.gpu

TOL = 1e-10
METRIC = ["cvalue", "Detmin", "invJ11", "invJ12", "invJ13", "invJ21", "invJ22", "invJ23",
          "dxdxi11", "dxdxi12", "dxdxi21", "dxdxi22", "J13dxi", "J23dxi", "J33dxi", "J13det",
          "J23det", "J33det"]
# the stored face averages (MeanKernel)
METRIC += [f"{q}_avg{d}" for d in ("EW", "SN")
           for q in ("invJ11", "invJ12", "invJ13", "invJ21", "invJ22", "invJ23", "Detmin")]
METRIC += ["dxdxi11_avgEW", "dxdxi21_avgEW", "dxdxi12_avgSN", "dxdxi22_avgSN"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fsum_mass(u, det):
    return (math.fsum((interior(u[0]) * interior(det)).ravel()),
            math.fsum((interior(u[3]) * interior(det)).ravel()))


@pytest.mark.parametrize("case", ["plane48", "hsiaolin", "df004"])
def test_metric_parity(case):
    sc = {"plane48": lambda: S.landslide_plane(n=48), "hsiaolin": hsiaolin, "df004": df004}[case]()
    say("metric parity", case, sc.nx, sc.ny)
    o = Oracle(sc)
    with Simulation(sc) as g:
        for name in METRIC:
            rel, ndiff = compare(g.field(name), o.get(name), name)
            assert ndiff == 0, (case, name, rel, ndiff)
        rel, ndiff = compare(g.field("svec"), o.get("svec"), "svec")
        assert ndiff == 0
        # the initial state (makeTopo10/11) including the ghost cells
        rel, ndiff = compare(g.state(), o.get("u"), "u0")
        assert ndiff == 0, (rel, ndiff)


@pytest.mark.parametrize("case", ["plane48", "hsiaolin", "df004", "fractal64"])
def test_uniform_metric_form_is_taken(case):
    """On the reference's uniform grids the step forms the inverse Jacobian from svec and
    cvalue instead of loading it (in FP64 its set-up check found the formed values equal to
    the stored ones bit for bit; every parity test below runs that form); so does the FP32
    variant, in float."""
    sc = {"plane48": lambda: S.landslide_plane(n=48), "hsiaolin": hsiaolin, "df004": df004,
          "fractal64": lambda: S.landslide_fractal(n=64, seed=3)}[case]()
    with Simulation(sc) as g:
        assert g.field("metric_uniform")[0, 0] == 1.0
    with Simulation(sc, precision="fp32") as g:
        assert g.field("metric_uniform")[0, 0] == 1.0


def run_pair(sc, steps, chunks, label, **kw):
    o = Oracle(sc)
    det = o.get("Detmin")
    with Simulation(sc, **kw) as g:
        done = 0
        for n in chunks:
            o.step(n)
            g.step(n)
            done += n
            ug, uo = g.state(), o.get("u")
            rel, ndiff = compare(interior(ug), interior(uo))
            dtg, dmaxg, _ = g.scalars()
            say(f"{label}: step {done} rel {rel:.3e} differing {ndiff} dt {dtg!r} / {o.scalar('dt')!r}")
            assert dtg == o.scalar("dt"), (done, dtg, o.scalar("dt"))
            assert dmaxg == o.scalar("dmax")
            assert rel <= TOL, (label, done, rel, ndiff)
            assert np.isfinite(ug).all()
        ms, mf = g.mass()
        rs, rf = fsum_mass(o.get("u"), det)
        assert abs(ms - rs) <= 1e-12 * abs(rs) and abs(mf - rf) <= 1e-12 * abs(rf), (ms, rs, mf, rf)
        tg, ntg = g.time_steps()
        assert ntg == done
        assert tg == o.scalar("time")
    return ndiff


@pytest.mark.parametrize("case,chunks", [("plane48", [1, 9, 30]), ("hsiaolin", [1, 4, 25]),
                                         ("df004", [1, 9, 50]), ("fractal64", [2, 28])])
def test_stored_metric_form_matches_oracle(case, chunks):
    """metric_form = 1: the step loads the stored inverse Jacobian and coordinate-derivative
    averages, as it does on any grid that fails the uniform-metric check.  That path gives
    the oracle's states, dt and mass as well."""
    sc = {"plane48": lambda: S.landslide_plane(n=48), "hsiaolin": hsiaolin, "df004": df004,
          "fractal64": lambda: S.landslide_fractal(n=64, seed=5)}[case]()
    with Simulation(sc, metric_form=1) as g:
        assert g.field("metric_uniform")[0, 0] == 0.0
    run_pair(sc, sum(chunks), chunks, case + "-stored-metric", metric_form=1)


def test_metric_forms_agree_bit_for_bit():
    """FP64 on a 160x96 fractal DEM: the formed metric (with dry-tile skipping) and the
    loaded one (every tile computed) give the same states, dt and times after 150 steps."""
    topo = S.fractal_dem(160, 96, 400.0, 0.8, 11)
    depth = S.paraboloid(160, 96, 20.0, 1600.0, 960.0, 1.0e6, 300.0)
    sc = S.Scenario(mode=0, topo=topo, depth=depth, cellsize=20.0, phiS0=0.5, delta0=20.0)
    with Simulation(sc) as a, Simulation(sc, metric_form=1, dry_skip=False) as b:
        assert a.field("metric_uniform")[0, 0] == 1.0 and b.field("metric_uniform")[0, 0] == 0.0
        for n in (1, 49, 100):
            a.step(n)
            b.step(n)
            assert a.scalars() == b.scalars()
            assert np.array_equal(a.state(), b.state())
        assert a.time_steps() == b.time_steps()


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 5), (5, 3), (7, 70), (64, 9), (65, 66),
                                   (127, 4), (130, 129)])
def test_parity_odd_grid_shapes(nx, ny):
    """Grids narrower than a stage tile or CFL box, and widths and heights one past a tile
    or box multiple: partial tiles, boxes and row segments at every edge, a release that
    reaches the ghost ring."""
    dx = 20.0
    r = max(0.35 * min(nx, ny) * dx, 1.6 * dx)
    depth = S.paraboloid(nx, ny, dx, 0.5 * (nx - 1) * dx, 0.5 * (ny - 1) * dx,
                         10.0 * math.pi * r * r / 2.0, r)
    topo = S.inclined_plane(nx, ny, dx, angle_deg=25.0, x_break=0.6 * nx * dx,
                            blend=0.2 * nx * dx) + S.gaussian_bump(nx, ny, dx, 0.3 * nx * dx,
                                                                    0.7 * ny * dx, 8.0, 3 * dx)
    sc = S.Scenario(mode=0, topo=topo, depth=depth, cellsize=dx, phiS0=0.55, delta0=22.0)
    run_pair(sc, 25, [1, 4, 20], f"odd{nx}x{ny}")


def test_parity_plane_small():
    sc = S.landslide_plane(n=48)
    assert run_pair(sc, 60, [1, 1, 3, 15, 40], "plane48") == 0


def test_parity_plane_c1_knobs():
    # the 256^2 configuration's knobs (CFL 0.4, density ratio 1000/2700) on a 96^2 grid
    sc = S.landslide_plane(n=96)
    sc.cfl = 0.4
    sc.rho_ratio = 1000.0 / 2700.0
    run_pair(sc, 80, [1, 9, 70], "plane96-cfl0.4")


def test_parity_hsiaolin():
    sc = hsiaolin()
    run_pair(sc, 40, [1, 4, 35], "hsiaolin")


def test_parity_df004_inflow():
    sc = df004()
    run_pair(sc, 120, [1, 9, 110], "df004")


def test_parity_fractal_per_cell_phi():
    sc = S.landslide_fractal(n=64, seed=7)
    run_pair(sc, 40, [2, 38], "fractal64-phi")


def test_parity_set_state_midrun():
    """Both implementations restarted from the same perturbed mid-run state."""
    sc = S.landslide_plane(n=40)
    o = Oracle(sc)
    o.step(10)
    u = o.get("u").copy()
    rng = np.random.default_rng(5)
    u[1] += 0.01 * rng.standard_normal(u[1].shape) * (u[0] > 0)
    u[5] -= 0.01 * rng.standard_normal(u[5].shape) * (u[3] > 0)
    o.set_u(u)
    with Simulation(sc) as g:
        g.set_state(u)
        o.step(10)
        g.step(10)
        rel, ndiff = compare(interior(g.state()), interior(o.get("u")))
        say("set_state midrun", rel, ndiff)
        assert rel <= TOL


def test_run_to_matches_oracle_clipping():
    """m2p_run_to(t_end): the step that reaches t_end is shortened to end on it (the
    oracle's output-time clipping with tf = [0, t_end]) and nothing runs after it."""
    sc = S.landslide_plane(n=48)
    o0 = Oracle(sc)
    o0.step(7)
    t_end = 0.5 * (o0.scalar("time") + o0.scalar("time") - o0.scalar("dt"))  # inside step 7
    o = Oracle(sc, tf=np.array([0.0, t_end, 1e30]))   # (a last entry it never reaches)
    o.step(7)
    assert o.scalar("time") == t_end
    with Simulation(sc) as g:
        n = g.run_to(t_end, max_steps=1000)
        tg, ntg = g.time_steps()
        assert (n, ntg, tg) == (7, 7, t_end)
        rel, ndiff = compare(interior(g.state()), interior(o.get("u")))
        assert ndiff == 0, (rel, ndiff)
        g.step(3)                 # plain stepping resumes after run_to
        o.step(3)
        rel, ndiff = compare(interior(g.state()), interior(o.get("u")))
        assert ndiff == 0 and g.time_steps()[1] == 10
        assert g.run_to(t_end) == 0   # already past it


def test_c1_workload_parity_256():
    """The 256^2 landslide configuration (35 degree plane, bump, 1e6 m^3 release,
    phi 0.6, CFL 0.4, density ratio 1000/2700) for 100 steps."""
    sc = S.landslide_plane(n=256, phi=0.6)
    sc.cfl = 0.4
    sc.rho_ratio = 1000.0 / 2700.0
    run_pair(sc, 100, [100], "c1-256")


def _threads():
    from oracle import oracle as O
    import os
    return O.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))


def test_c1_workload_parity_full_2000_steps():
    """The 256^2 landslide configuration for its full 2000 steps, bit for bit (the oracle
    runs its loops on all host threads; results do not depend on the thread count)."""
    say("oracle threads", _threads())
    sc = S.landslide_plane(n=256, phi=0.6)
    sc.cfl = 0.4
    sc.rho_ratio = 1000.0 / 2700.0
    ndiff = run_pair(sc, 2000, [500, 500, 1000], "c1-256-full")
    assert ndiff == 0


def test_c3_inflow_valley_full_size():
    """The 1024 x 4096 inflow configuration (V-shaped valley, triangular hydrograph on the
    upstream edge): 30 steps, bit for bit."""
    _threads()
    sc = S.inflow_valley()
    assert (sc.nx, sc.ny) == (1024, 4096)
    ndiff = run_pair(sc, 30, [1, 29], "c3-inflow")
    assert ndiff == 0


def test_c2_landslide_2048_full_size():
    """The 2048^2 throughput configuration: 10 steps, bit for bit."""
    _threads()
    sc = S.landslide_fractal(n=2048, seed=2)
    ndiff = run_pair(sc, 10, [2, 8], "c2-2048")
    assert ndiff == 0


def test_gpu_matches_committed_golden_states():
    """The HIP path against the oracle states frozen in tests/golden/oracle_golden.npz."""
    import os
    from util import GOLDEN
    sys.path.insert(0, GOLDEN)
    from make_golden import CASES, digest
    gold = np.load(os.path.join(GOLDEN, "oracle_golden.npz"))
    for name, (make, steps) in CASES.items():
        with Simulation(make()) as g:
            g.step(steps)
            u = np.ascontiguousarray(interior(g.state()))
            t, _ = g.time_steps()
        h = digest(u)
        say("golden", name, h[:16], str(gold[f"{name}/sha256"])[:16])
        assert h == str(gold[f"{name}/sha256"]), name
        assert t == float(gold[f"{name}/time"])
