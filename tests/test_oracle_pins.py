"""Pins of the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, expected outputs or golden vectors (its data files are
inputs only), so each part of the oracle is pinned here by a closed form, an invariant,
a textbook special case or brute force, chosen so that a plausible transcription mistake
(a dropped term, a wrong sign, a swapped index or operand) fails one of them.
"""
import math

import numpy as np
import pytest

from moses_2pdf_amd import scenarios as S
from oracle.oracle import Oracle, f32
from util import GOLDEN, df004, hsiaolin

H0 = 1.0e-6


def flat_lake(n=24, depth=2.0, phi=0.5):
    topo = np.zeros((n, n))
    d = np.full((n, n), depth)
    return S.Scenario(mode=0, topo=topo, depth=d, cellsize=20.0, phiS0=phi, delta0=20.0)


def mass(o):
    u, D = o.get("u"), o.get("Detmin")
    return (math.fsum((u[0, 3:-3, 3:-3] * D[3:-3, 3:-3]).ravel()),
            math.fsum((u[3, 3:-3, 3:-3] * D[3:-3, 3:-3]).ravel()))


# ---------------------------------------------------------------------------------------
# terrain metric (makeTopo1..11, JacobKernel, MeanKernel)
# ---------------------------------------------------------------------------------------

def test_metric_on_inclined_plane_closed_forms():
    # z = 1000 - 10*col metres: exact in float, slope -0.5 along x.  The edge copy of the
    # topography into the ghosts halves the slope on the first/last interior column, so
    # the slope-dependent closed forms are checked on columns 4..N-5.
    n, dx, t = 32, 20.0, 0.5
    topo = (1000.0 - 10.0 * np.arange(n)[None, :]).repeat(n, axis=0)
    sc = S.Scenario(mode=0, topo=topo, depth=np.zeros((n, n)), cellsize=dx)
    o = Oracle(sc)
    I = (slice(3, -3), slice(4, -4))
    bx = o.get("dbdx")[I]
    assert np.allclose(bx, -t, rtol=1e-15)
    assert np.allclose(o.get("dbdx")[3:-3, 3], -t / 2, rtol=1e-15)  # one-sided at the edge
    assert np.all(o.get("dbdy")[3:-3, 3:-3] == 0.0)
    c = o.get("cvalue")[I]
    assert np.allclose(c, 1.0 / math.sqrt(1.0 + t * t), rtol=1e-15)
    s = o.get("svec")
    assert np.allclose(s[0][I], c * bx, rtol=1e-15) and np.all(s[1][I] == 0.0)
    # Cartesian grid: dx/dxi = I; Detmin = 1/c; invJ of the surface frame
    assert np.allclose(o.get("dxdxi11")[I], 1.0, rtol=1e-15)
    assert np.all(o.get("dxdxi12")[I] == 0.0) and np.all(o.get("dxdxi21")[I] == 0.0)
    assert np.allclose(o.get("Detmin")[I], 1.0 / c, rtol=1e-14)
    sin2 = (c * bx) ** 2
    assert np.allclose(o.get("invJ11")[I], 1.0 - sin2, rtol=1e-14)
    assert np.allclose(o.get("invJ13")[I], c * (c * bx), rtol=1e-14)
    assert np.allclose(o.get("invJ22")[I], 1.0, rtol=1e-15)
    assert np.all(o.get("invJ12")[I] == 0.0) and np.all(o.get("invJ21")[I] == 0.0)
    assert np.allclose(o.get("invJ31")[I], -(c * bx), rtol=1e-15)
    assert np.allclose(o.get("invJ33")[I], c, rtol=1e-15)
    I = (slice(3, -3), slice(5, -5))
    # a plane has no curvature
    for nm in ("J13dxi", "J23dxi", "J33dxi", "J13det", "J23det", "J33det"):
        assert np.abs(o.get(nm)[I]).max() < 1e-12, nm
    # ghost copies: column 0 equals column 3, row 0 equals row 3
    cv = o.get("cvalue")
    assert np.all(cv[:, 0] == cv[:, 3]) and np.all(cv[0, :] == cv[3, :])
    # the reference's dxdxi21 quirk: column nx+MD+2 is never written by Boundary1Kernel
    assert np.all(o.get("dxdxi21")[:, -1] == 0.0)


def test_curvature_terms_on_a_parabolic_valley():
    """b = k x^2 / 2 across x: J13dxi = d(-s1)/dx = -k c^3 and J33dxi = dc/dx = -k^2 x c^3,
    with second-order central differences (relative error O(dx^2))."""
    n, dx, k = 48, 20.0, 2.0e-4
    x = (np.arange(n) - n // 2) * dx
    topo = np.repeat((0.5 * k * x * x)[None, :], n, axis=0) + 100.0
    sc = S.Scenario(mode=0, topo=topo.astype(np.float32).astype(np.float64),
                    depth=np.zeros((n, n)), cellsize=dx)
    o = Oracle(sc)
    K = k / 10.0  # scaled lengths (x10): b and x both scale, curvature scales by 1/10
    X = (np.arange(n + 6) - 3 - n // 2) * dx * 10.0
    c = 1.0 / np.sqrt(1.0 + (K * X) ** 2)
    cols = slice(8, n - 2)
    row = n // 2
    j13 = o.get("J13dxi")[row, cols]
    j33 = o.get("J33dxi")[row, cols]
    assert np.allclose(j13, (-K * c ** 3)[cols], rtol=2e-3)
    assert np.allclose(j33, (-K * K * X * c ** 3)[cols], rtol=2e-3, atol=1e-12)
    assert np.abs(o.get("J13det")[3:-3, 3:-3]).max() < 1e-15  # nothing varies along y


# ---------------------------------------------------------------------------------------
# TVD slopes (TVD1..5Kernel)
# ---------------------------------------------------------------------------------------

def minmod(a, b):
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


def test_slopes_are_textbook_minmod():
    sc = flat_lake(n=20)
    o = Oracle(sc)
    rng = np.random.default_rng(3)
    u = o.get("u").copy()
    for k in range(7):
        u[k] = rng.standard_normal(u[k].shape)
        u[k][rng.random(u[k].shape) < 0.2] = 0.0      # zero differences: sign(0) = -1
        u[k][:, 7] = u[k][:, 8]                         # equal neighbours
    o.set_u(u)
    o.faces()
    ug = o.get("u")   # after the ghost fill
    ux, uy = o.get("ux"), o.get("uy")
    dA = ug[:, 1:-1, 1:-1] - ug[:, 1:-1, :-2]
    dB = ug[:, 1:-1, 2:] - ug[:, 1:-1, 1:-1]
    assert np.array_equal(ux[:, 1:-1, 1:-1], minmod(dA, dB))
    dA = ug[:, 1:-1, 1:-1] - ug[:, :-2, 1:-1]
    dB = ug[:, 2:, 1:-1] - ug[:, 1:-1, 1:-1]
    assert np.array_equal(uy[:, 1:-1, 1:-1], minmod(dA, dB))
    # Boundary9Kernel leaves the outer ring of slopes at zero
    assert np.all(ux[:, 0, :] == 0) and np.all(ux[:, :, -1] == 0)


# ---------------------------------------------------------------------------------------
# interface states, positivity, fluxes and update
# ---------------------------------------------------------------------------------------

def test_lake_at_rest_stays_at_rest():
    """Flat ground, uniform layer, no motion: depths and momenta never change, the
    hydrostatic fluxes cancel exactly; phi is recomputed as h_s/(h_s+h_f+H0) by the first
    update and then stays fixed."""
    o = Oracle(flat_lake(n=24, depth=3.0))
    u0 = o.get("u")[:, 3:-3, 3:-3].copy()
    o.step(1)
    phi1 = o.get("u")[6, 3:-3, 3:-3].copy()
    o.step(29)
    u1 = o.get("u")[:, 3:-3, 3:-3]
    assert np.array_equal(u0[:6], u1[:6])
    assert np.all(u1[1] == 0) and np.all(u1[2] == 0) and np.all(u1[4] == 0) and np.all(u1[5] == 0)
    assert np.all(phi1 == u0[0] / (u0[0] + u0[3] + H0)) and np.array_equal(u1[6], phi1)


def test_interface_states_of_linear_data_are_exact():
    """For data linear along x, minmod gives the exact slope and uE, uW at each face are the
    face value (u[c] + u[c+1])/2."""
    sc = flat_lake(n=20)
    o = Oracle(sc)
    u = o.get("u").copy()
    cols = np.arange(u.shape[2], dtype=np.float64)
    for k in range(7):
        u[k] = (1.0 + k) + 0.25 * cols[None, :]
    o.set_u(u)
    o.faces()
    uE, uW = o.get("uE"), o.get("uW")
    # faces 3..nx+1 (the right boundary face uses the extrapolated ghost cell nx+3)
    rows, cs = slice(3, -3), slice(3, -5)
    face = 1.0 + 0.25 * (cols[3:-5] + 0.5)
    for k in (1, 2, 4, 5):  # momenta (h stays positive, phi is recomputed)
        assert np.array_equal(uE[k][rows, cs], np.broadcast_to(k + face, uE[k][rows, cs].shape))
        assert np.array_equal(uW[k][rows, cs], np.broadcast_to(k + face, uW[k][rows, cs].shape))


def test_mass_is_conserved():
    sc = S.landslide_plane(n=64)
    o = Oracle(sc)
    m0 = mass(o)
    o.step(60)
    m1 = mass(o)
    assert abs(m1[0] - m0[0]) <= 1e-13 * m0[0]
    assert abs(m1[1] - m0[1]) <= 1e-13 * m0[1]


def test_mirror_symmetry_across_the_slope_axis():
    n, dx = 64, 20.0
    s = n * dx / 5120.0
    topo = S.inclined_plane(n, n, dx, x_break=3000 * s, blend=250 * s)
    depth = S.paraboloid(n, n, dx, 1000 * s, (n - 1) * dx / 2, 1e6 * s * s, 150 * s)
    sc = S.Scenario(mode=0, topo=topo, depth=depth, cellsize=dx, phiS0=0.6, delta0=20.0)
    o = Oracle(sc)
    o.step(40)
    u = o.get("u")[:, 3:-3, 3:-3]
    for k in (0, 1, 3, 4, 6):
        assert np.array_equal(u[k], u[k][::-1]), k
    for k in (2, 5):
        assert np.array_equal(u[k], -u[k][::-1]), k


# ---------------------------------------------------------------------------------------
# dt: reduceKernel (float-rounded max), CFL3Kernel clamp
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("depth,expect", [(50.0, "free"), (1e-12, "cap"), (1e8, "floor")])
def test_dt_closed_form_for_a_lake(depth, expect):
    phi, cfl = 0.5, 0.1
    o = Oracle(flat_lake(n=16, depth=depth, phi=phi))
    o.faces()
    H = 10.0 * depth                      # scaled total depth (phi*c + (1-phi)*c, c = 1)
    hs = H * phi
    phif = hs / (hs + (H - hs) + H0)      # KeepPositivi2Kernel
    dx = 200.0
    em_f = math.sqrt(H / (1.0 - phif + H0))   # fluid speed, q = 0, invJ11 = 1
    em_s = math.sqrt(H * (1.0 - 1.420 / 2.60))
    em = max(em_f, em_s, 1e-10)
    dmax = float(np.float32(em / dx))
    assert o.scalar("dmax") == dmax
    dt = min(2.0, cfl / dmax)
    dt = max(0.001, dt)
    assert o.scalar("dtval") == dt
    assert {"free": 0.001 < dt < 2.0, "cap": dt == 2.0, "floor": dt == 0.001}[expect]


def test_output_times_clip_dt():
    sc = S.landslide_plane(n=24)
    sc.total_sim, sc.each_out = 3.0, 1.0
    o = Oracle(sc, outputs=True)
    hits = []
    for _ in range(200):
        if o.step(1) == 0:
            break
        if o.scalar("lastWritten"):
            hits.append((int(o.scalar("lastWritten")), o.scalar("time")))
    assert [h[0] for h in hits] == [2, 3, 4]
    for it, t in hits:
        assert abs(t - (it - 1) * 10.0) <= 1e-12
    assert o.scalar("nstop") == 1


# ---------------------------------------------------------------------------------------
# velocity-crossing limiter (Flux22..28)
# ---------------------------------------------------------------------------------------

def test_crossing_limiter_prevents_relative_velocity_reversal():
    sc = flat_lake(n=20, depth=1.0)
    sc.Cd = 4000.0           # strong drag: the explicit update overshoots
    o = Oracle(sc)
    u = o.get("u").copy()
    us, uf = 0.8, -0.8
    u[1] = u[0] * us
    u[4] = u[3] * uf
    o.set_u(u)
    o.step(1)
    one, old = o.get("uone")[:, 3:-3, 3:-3], o.get("uzero")[:, 3:-3, 3:-3]
    s_new = one[1] / (one[0] + H0)
    f_new = one[4] / (one[3] + H0)
    d_old = old[1] / (old[0] + H0) - old[4] / (old[3] + H0)
    fired = np.abs(s_new - f_new) <= np.abs(s_new) * H0 / one[3] * 1.01
    assert fired.any()                      # the drag overshoot did flip the sign
    assert np.all((s_new - f_new) * d_old >= -np.abs(d_old * s_new) * 1e-5)


# ---------------------------------------------------------------------------------------
# inflow injection (Inflow1/2Kernel, the dire rule)
# ---------------------------------------------------------------------------------------

def test_df004_injection_strips():
    sc = df004(phi=0.5)
    o = Oracle(sc)
    before = o.get("u").copy()
    o.inject()
    u = o.get("u")
    c = o.get("cvalue")
    changed = np.argwhere((u != before).any(axis=0))
    # locations X = 43, Y = 13..20 -> 6x1 strips at columns 43..48 (padded) on rows 13..20
    expect = {(r, col) for r in range(13, 21) for col in range(43, 49)}
    assert {tuple(x) for x in changed} <= expect
    H = 10.0 * f32(0.017)
    for r in range(13, 21):
        for col in range(43, 49):
            assert u[0, r, col] == H * (1 - 0.5) / 0.5
            assert u[3, r, col] == H * (1.0 - 0.5) * c[r, col]
            assert u[6, r, col] == u[0, r, col] / (u[0, r, col] + u[3, r, col] + H0)


# ---------------------------------------------------------------------------------------
# self-generated golden states (tests/golden/make_golden.py)
# ---------------------------------------------------------------------------------------

def test_oracle_reproduces_golden_states():
    import os
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import CASES, digest, run
    g = np.load(os.path.join(GOLDEN, "oracle_golden.npz"))
    for name in CASES:
        u, t, dt = run(name)
        h = digest(u)
        assert h == str(g[f"{name}/sha256"]), name
        assert t == float(g[f"{name}/time"]) and dt == float(g[f"{name}/dt"])


def test_landslide_ghost_momenta_stay_zero():
    # Boundary7/8 (L/K:905-956) give a landslide's momenta ghost cells 0 in every stage and
    # nothing else writes a landslide's ghost cells, so after any step they hold 0; the
    # HIP ghost fills rely on this and leave those planes alone (an inflow run's strips
    # write momenta into ghost cells, so inflow fills all planes)
    sc = S.landslide_fractal(n=48, seed=3)
    o = Oracle(sc)
    ring = np.ones((o.nyd, o.nxd), dtype=bool)
    ring[3:-3, 3:-3] = False
    moved = False
    for _ in range(4):
        o.step(5)
        u = o.get("u")
        for k in (1, 2, 4, 5):
            assert not np.any(u[k][ring]), k
        moved = moved or bool(np.any(u[1][~ring] != 0))
    assert moved


def test_dry_cells_hold_positive_zero_in_every_component():
    # Flux20 (L/K:2682-2709) sets a dry phase's depth and momenta to +0 and Flux21 gives
    # phi = 0/(0+0+H0) = +0, so after a step every interior cell without material holds +0
    # bit for bit in all 7 components: a tile that skips such cells leaves exactly what
    # computing them writes
    sc = S.landslide_fractal(n=48, seed=3)
    o = Oracle(sc)
    o.step(12)
    u = o.get("u")[:, 3:-3, 3:-3]
    dry = (u[0] == 0) & (u[3] == 0)
    assert dry.sum() > 100 and (~dry).sum() > 10
    bits = u.view(np.uint64)
    for k in range(7):
        assert not np.any(bits[k][dry]), k


@pytest.mark.parametrize("slope,v0,rel", [(0.5, 0.0, 1e-7), (0.1, 3.0, 5e-5)])
def test_sliding_layer_on_an_incline_closed_form(slope, v0, rel):
    # A uniform solid layer (phiS0 = 1, no fluid) on a plane of slope tan(theta): every
    # flux difference vanishes in the interior, so the horizontal velocity us = u1/u0
    # follows the sources alone, and a block sliding with Coulomb friction on an incline
    # accelerates along the slope at g (sin(theta) - mu cos(theta)), horizontally
    # c^2 (tan(theta) - mu) with c = cos(theta), g = 1 in the scaled units and
    # mu = (1 - rho_r) tan(delta) (the basal load is buoyancy-reduced, Flux16/17
    # L/K:2530-2621; tan(delta) with the reference's 3.14159, L/KH:17).  The RK2 steps
    # integrate a constant acceleration exactly once the layer moves.  Sliding down a
    # 0.5 slope from rest, and braking on a 0.1 slope from us = 3, pin the sign and size
    # of gravity and friction.  Tolerance: the drag term with no fluid is
    # Cd*phi*(1-phi)*h*us ~ Cd*H0*us, about 1e-5 of the acceleration at us = 3.
    n, dx = 48, 20.0
    x = np.arange(n) * dx
    topo = np.tile((n * dx - x) * slope, (n, 1))   # exact in float after the x10 scaling
    sc = S.Scenario(mode=0, topo=topo, depth=np.full((n, n), 2.0), cellsize=dx, phiS0=1.0,
                    delta0=20.0)
    o = Oracle(sc)
    if v0:
        u = o.get("u")
        u[1] = u[0] * v0
        o.set_u(u)
    o.step(2)   # the first step switches the friction on
    c = 3 + n // 2
    u_a, t_a = o.get("u"), o.scalar("time")
    o.step(6)
    u_b, t_b = o.get("u"), o.scalar("time")
    acc = (u_b[1][c, c] / u_b[0][c, c] - u_a[1][c, c] / u_a[0][c, c]) / (t_b - t_a)
    mu = (1.0 - sc.rho_ratio) * math.tan(20.0 * 3.14159 / 180.0)
    expect = (slope - mu) / (1.0 + slope * slope)
    assert acc == pytest.approx(expect, rel=rel)
    assert u_b[0][c, c] == u_a[0][c, c]   # the depth does not change
    assert u_b[2][c, c] == 0.0            # nor does the cross-slope momentum


@pytest.mark.parametrize("slope,v0", [(0.3, 0.0), (0.0, 5.0)])
def test_fluid_layer_on_an_incline_follows_the_midpoint_rule(slope, v0):
    # A uniform fluid layer (phiS0 = 0, no solid) on a plane of slope tan(theta): the flux
    # differences vanish in the interior and the fluid feels gravity on its load and the
    # linear resistance of Flux17 (L/K:2562-2621), so its horizontal velocity obeys
    #   dv/dt = r (c^2 tan(theta) - (varTheta / N_R) v),   r = hf / (hf + H0),
    # with v = u4 / (u3 + H0) as the scheme forms it (r is that regularisation).  The
    # predictor (dt/2) and corrector (dt) are the explicit midpoint rule, so v follows the
    # midpoint recurrence of this ODE step by step, with each step's dt read back, to
    # round-off.  Accelerating from rest on a 0.3 slope and relaxing from v = 5 on flat
    # ground pin the gravity term and the resistance (its sign and varTheta / N_R).
    n, dx = 48, 20.0
    x = np.arange(n) * dx
    topo = np.tile((n * dx - x) * slope, (n, 1))
    sc = S.Scenario(mode=0, topo=topo, depth=np.full((n, n), 2.0), cellsize=dx, phiS0=0.0,
                    delta0=20.0)
    o = Oracle(sc)
    c = 3 + n // 2
    if v0:
        u = o.get("u")
        u[4] = u[3] * v0
        o.set_u(u)
    H0 = 1.0e-6
    A, k = slope / (1.0 + slope * slope), sc.varTheta / sc.N_R
    u = o.get("u")
    r = u[3][c, c] / (u[3][c, c] + H0)
    v, t = u[4][c, c] / (u[3][c, c] + H0), o.scalar("time")
    f = lambda w: r * (A - k * w)
    for _ in range(15):
        o.step(1)
        t_next = o.scalar("time")
        dt = t_next - t
        v = v + dt * f(v + 0.5 * dt * f(v))
        t = t_next
        u = o.get("u")
        assert u[4][c, c] / (u[3][c, c] + H0) == pytest.approx(v, rel=1e-12)
        assert u[3][c, c] / (u[3][c, c] + H0) == r   # the depth does not change
    assert abs(v - v0) > 0.5   # the velocity did move


def test_phase_drag_conserves_momentum_and_relaxes():
    # On flat ground with no basal friction (delta0 = 0) and no fluid resistance
    # (varTheta = 0), the only source in a uniform two-phase layer is the drag between the
    # phases (Flux17 L/K:2562-2621), which only moves momentum from one phase to the
    # other: the density-weighted total u1 + rho_r u4 (solid plus fluid momentum over the
    # solid density) must stay constant step after step, and the relative velocity must
    # decay without changing sign (Cd small enough that the explicit step does not
    # overshoot, so the velocity-crossing limiter stays idle).
    n, dx = 48, 20.0
    sc = S.Scenario(mode=0, topo=np.zeros((n, n)), depth=np.full((n, n), 2.0), cellsize=dx,
                    phiS0=0.5, delta0=0.0, Cd=0.5, varTheta=0.0)
    o = Oracle(sc)
    c = 3 + n // 2
    u = o.get("u")
    u[4] = u[3] * 2.0   # the fluid moves, the solid rests
    o.set_u(u)
    H0, rr = 1.0e-6, sc.rho_ratio
    Q0 = u[1][c, c] + rr * u[4][c, c]
    rel = [2.0]
    for _ in range(20):
        o.step(1)
        u = o.get("u")
        assert u[1][c, c] + rr * u[4][c, c] == pytest.approx(Q0, rel=1e-13)
        rel.append(u[4][c, c] / (u[3][c, c] + H0) - u[1][c, c] / (u[0][c, c] + H0))
    assert all(a > b > 0.0 for a, b in zip(rel, rel[1:]))
    assert rel[-1] < 1e-4 * rel[0]


def _flat_solid_layer_after_a_step():
    # a random solid-only state (phiS0 = 1) on flat ground, one step on: the face arrays
    # then all belong to the corrector stage
    n, dx = 32, 20.0
    rng = np.random.default_rng(7)
    sc = S.Scenario(mode=0, topo=np.zeros((n, n)), depth=1.0 + 0.5 * rng.random((n, n)),
                    cellsize=dx, phiS0=1.0, delta0=20.0)
    o = Oracle(sc)
    u = o.get("u")
    u[1] = u[0] * rng.uniform(-1, 1, u[0].shape)
    u[2] = u[0] * rng.uniform(-1, 1, u[0].shape)
    o.set_u(u)
    o.step(1)
    return sc, o


def test_depth_flux_is_the_central_upwind_combination():
    # On flat ground the metric is the identity (Det = 1, invJ = I, no surface slope), so
    # the physical flux of the solid depth through an x-face is h u with u = m / (h + H0),
    # and the numerical flux (Flux8, L/K:2039-2211) must be the central-upwind / Rusanov
    # combination of the two face states with the face's wave speed a:
    #   H = 1/2 (F_E + F_W - a Det (h_E - h_W)),  E = right state, W = left state
    # (likewise for y-faces with N/S), on every interior face.
    sc, o = _flat_solid_layer_after_a_step()
    H0 = 1.0e-6
    inner = (slice(4, -4), slice(4, -4))
    for a, b, k, ap, det, H in (("uE", "uW", 1, "apEW", "Detmin_avgEW", "Hpx"),
                                ("uN", "uS", 2, "apSN", "Detmin_avgSN", "Hpy")):
        ua, ub, c, D = o.get(a), o.get(b), o.get(ap), o.get(det)
        assert np.all(D[inner] == 1.0)
        Fa = ua[0] * (ua[k] / (ua[0] + H0)) * D
        Fb = ub[0] * (ub[k] / (ub[0] + H0)) * D
        expect = 0.5 * (Fa + Fb - c * D * (ua[0] - ub[0]))
        got = o.get(H)[0]
        assert np.abs(c * D * (ua[0] - ub[0]))[inner].mean() > 0.1 * np.abs(got[inner]).mean()
        np.testing.assert_allclose(got[inner], expect[inner], rtol=1e-13, atol=1e-13)


def test_momentum_flux_and_wave_speed_are_shallow_water():
    # On flat ground a solid-only layer is a shallow-water layer under the buoyancy-reduced
    # gravity g' = 1 - rho_r (g = 1): the flux of the normal momentum m through a face is
    # m u + g' h^2 / 2 (Flux6/7 L/K:1773-1936), combined central-upwind as the depth flux,
    # and the face's wave speed is the larger of |u| + sqrt(g' h) over its two states
    # (Flux4/5 L/K:1685-1770).
    sc, o = _flat_solid_layer_after_a_step()
    H0, g = 1.0e-6, 1.0 - sc.rho_ratio
    inner = (slice(4, -4), slice(4, -4))
    for a, b, k, ap, H in (("uE", "uW", 1, "apEW", "Hpx"), ("uN", "uS", 2, "apSN", "Hpy")):
        ua, ub, c = o.get(a), o.get(b), o.get(ap)
        va, vb = ua[k] / (ua[0] + H0), ub[k] / (ub[0] + H0)
        speed = np.maximum(np.abs(va) + np.sqrt(g * ua[0]), np.abs(vb) + np.sqrt(g * ub[0]))
        np.testing.assert_allclose(c[inner], speed[inner], rtol=1e-14)
        Fa = ua[k] * va + 0.5 * g * ua[0] ** 2
        Fb = ub[k] * vb + 0.5 * g * ub[0] ** 2
        expect = 0.5 * (Fa + Fb - c * (ua[k] - ub[k]))
        np.testing.assert_allclose(o.get(H)[k][inner], expect[inner], rtol=1e-13, atol=1e-13)


def test_fluid_fluxes_are_shallow_water():
    # The fluid branch of the same fluxes: a fluid-only layer (phiS0 = 0) on flat ground is
    # a shallow-water layer under the full gravity g = 1, whose fluid depth and normal
    # momentum fluxes are the central-upwind combinations of h u and m u + h^2 / 2 with the
    # fluid wave speed |u| + sqrt(h / (1 - phi + H0)) (phi = 0 here) of each face.
    n, dx = 32, 20.0
    rng = np.random.default_rng(8)
    sc = S.Scenario(mode=0, topo=np.zeros((n, n)), depth=1.0 + 0.5 * rng.random((n, n)),
                    cellsize=dx, phiS0=0.0, delta0=20.0)
    o = Oracle(sc)
    u = o.get("u")
    u[4] = u[3] * rng.uniform(-1, 1, u[3].shape)
    u[5] = u[3] * rng.uniform(-1, 1, u[3].shape)
    o.set_u(u)
    o.step(1)
    H0 = 1.0e-6
    inner = (slice(4, -4), slice(4, -4))
    for a, b, k, H in (("uE", "uW", 4, "Hpx"), ("uN", "uS", 5, "Hpy")):
        ua, ub = o.get(a), o.get(b)
        va, vb = ua[k] / (ua[3] + H0), ub[k] / (ub[3] + H0)
        pa, pb = ua[0] / (ua[0] + ua[3] + H0), ub[0] / (ub[0] + ub[3] + H0)
        speed = np.maximum(np.abs(va) + np.sqrt((ua[0] + ua[3]) / (1 - pa + H0)),
                           np.abs(vb) + np.sqrt((ub[0] + ub[3]) / (1 - pb + H0)))
        got = o.get(H)
        depth = 0.5 * (ua[3] * va + ub[3] * vb - speed * (ua[3] - ub[3]))
        mom = 0.5 * (ua[k] * va + 0.5 * ua[3] ** 2 + ub[k] * vb + 0.5 * ub[3] ** 2
                     - speed * (ua[k] - ub[k]))
        np.testing.assert_allclose(got[3][inner], depth[inner], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(got[k][inner], mom[inner], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("slope", [0.0, 0.5])
def test_fluid_shear_viscosity_closed_form(slope):
    # On flat ground the fluid viscous fluxes (Flux8-10, L/K:2039-2347) are the
    # depth-integrated Newtonian stresses h tau with tau_xy = mu (du/dy + dv/dx),
    # mu = (1 - phi) / N_R.  A fluid-only layer of uniform depth with the shear profile
    # u = beta y^2, v = 0 (y counted from 100 cells below the grid, so the profile is
    # monotone and the minmod face-state jumps are all equal, which makes the upwind
    # dissipation cancel between faces) and no resistance (varTheta = 0) then gains
    # du/dt = mu d2u/dy2 = 2 beta / N_R everywhere (times h / (h + H0), the scheme's
    # regularisation): after time t every interior row has gained r 2 beta t / (dy^2 N_R).
    # On a plane tilted along x (slope 0.5) the y direction is untouched by the tilt, so
    # the transform of the gradients to surface coordinates (Flux9) must cancel its
    # tilt terms and leave the same viscous gain, on top of gravity's r c^2 tan(theta) t.
    n, dx = 40, 1.0
    x = np.arange(n) * dx
    sc = S.Scenario(mode=0, topo=np.tile((n * dx - x) * slope, (n, 1)),
                    depth=np.full((n, n), 2.0), cellsize=dx, phiS0=0.0, delta0=20.0,
                    varTheta=0.0, N_R=1.0)
    o = Oracle(sc)
    H0, beta = 1.0e-6, 1.0e-3
    u = o.get("u")
    q0 = beta * (np.arange(o.nyd)[:, None] + 100.0) ** 2 * np.ones((1, o.nxd))
    u[4] = (u[3] + H0) * q0
    o.set_u(u)
    c = o.nyd // 2
    r = u[3][c, c] / (u[3][c, c] + H0)
    dy = 10.0 * dx   # the reference's lengths are scaled by 10
    t0 = o.scalar("time")
    o.step(10)
    t = o.scalar("time") - t0
    u = o.get("u")
    band = slice(c - 4, c + 5)
    gained = u[4][band, c] / (u[3][band, c] + H0) - q0[band, c]
    visc = r * 2.0 * beta * t / (dy * dy * sc.N_R)
    gravity = r * slope / (1.0 + slope * slope) * t
    np.testing.assert_allclose(gained - gravity, visc, rtol=1e-7)
    assert np.all(np.abs(u[5][band, c]) < 1e-30)   # no cross-stream velocity appears


# ---------------------------------------------------------------------------------------
# viscous stress through the surface transform (Flux8-10) and the centripetal load
# (Flux15-16), read directly from one predictor's terms
# ---------------------------------------------------------------------------------------

def _film_stress_on_a_plane(tx, ty, a, b, c, d, h):
    """Depth-integrated Newtonian stress of a thin film on the plane z = tx x + ty y, from
    vector geometry alone: the film's 3-D velocity is V = (u, v, tx u + ty v) with the
    horizontal velocity u = a x + b y, v = c x + d y; its surface strain rate is the
    symmetric part of the tangential velocity gradient; the stress (per unit viscosity) is
    twice that; the force on a face is the depth h times the stress applied to the face's
    unit normal within the surface, times the face's length per unit grid coordinate.
    Returns the horizontal (x, y) components of the force on an x-face (a line x = const)
    and on a y-face."""
    ex, ey = np.array([1.0, 0.0, tx]), np.array([0.0, 1.0, ty])
    n = np.array([-tx, -ty, 1.0])
    n /= np.linalg.norm(n)
    dVx = np.array([a, c, tx * a + ty * c])          # dV/dx along the surface
    dVy = np.array([b, d, tx * b + ty * d])          # dV/dy along the surface
    E = np.stack([ex, ey])
    G = E @ E.T                                       # metric of the (x, y) parametrisation
    Lg = np.stack([dVx, dVy], axis=1) @ np.linalg.inv(G) @ E   # tangential gradient
    Pn = np.eye(3) - np.outer(n, n)
    tau = Pn @ (Lg + Lg.T) @ Pn                       # 2 x strain rate

    def force(t, toward):
        m = np.cross(t, n)
        m /= np.linalg.norm(m)
        if m @ toward < 0:
            m = -m
        return h * (tau @ m) * np.linalg.norm(t)
    fx, fy = force(ey, ex), force(ex, ey)
    return fx[0], fx[1], fy[0], fy[1]


@pytest.mark.parametrize("tx,ty", [(0.0, 0.0), (0.5, 0.0), (0.0, 0.5), (0.5, 0.25),
                                   (-0.25, 0.5)])
def test_viscous_fluxes_are_the_film_stress_on_a_plane(tx, ty):
    # The viscous fluxes of the fluid momenta (PDx4,5 through x-faces, PDy4,5 through
    # y-faces; Flux8-10, L/K:2039-2347) are the depth-integrated Newtonian stress of the
    # film, h tau with tau = 2 D (the viscosity (1 - phi)/N_R multiplies them in the
    # update), projected on the horizontal axes.  On a plane of any tilt a fluid layer of
    # uniform depth with a linear horizontal velocity field has a uniform strain rate, so
    # every interior face must carry the force computed from vector geometry above.  This
    # pins the factor 2 of the normal stresses (flat case), and every correction term of
    # the transform to surface coordinates (Flux9), each of which the tilted cases weigh
    # with a nonzero gradient.
    n, dx = 40, 1.0
    x = np.arange(n) * dx
    topo = 1000.0 + tx * x[None, :] + ty * x[:, None]      # exact in float
    sc = S.Scenario(mode=0, topo=topo, depth=np.full((n, n), 2.0), cellsize=dx, phiS0=0.0,
                    delta0=20.0, varTheta=0.0, N_R=1.0)
    o = Oracle(sc)
    u = o.get("u")
    u[3] = 20.0                    # uniform fluid depth (the edge columns' c differs)
    A, B, C, D = 1.0e-3, 2.0e-3, -3.0e-3, 5.0e-3            # increments per cell
    rows, cols = np.arange(o.nyd)[:, None], np.arange(o.nxd)[None, :]
    u[4] = (u[3] + H0) * (A * cols + B * rows + 0.1)
    u[5] = (u[3] + H0) * (C * cols + D * rows - 0.2)
    o.set_u(u)
    o.stage_terms()
    ds = 10.0 * dx                                          # lengths are scaled by 10
    want = _film_stress_on_a_plane(tx, ty, A / ds, B / ds, C / ds, D / ds, 20.0)
    PDx, PDy = o.get("PDx"), o.get("PDy")
    inner = (slice(5, -5), slice(5, -5))
    for got, w in ((PDx[4], want[0]), (PDx[5], want[1]), (PDy[4], want[2]), (PDy[5], want[3])):
        assert abs(w) > 1e-4
        np.testing.assert_allclose(got[inner], w, rtol=1e-11)
    for k in range(4):                                      # no viscous flux of depth or solid
        assert np.all(PDx[k] == 0.0) and np.all(PDy[k] == 0.0)


@pytest.mark.parametrize("k1,k12,k2", [(1 / 512, 0.0, 0.0), (0.0, 0.0, 1 / 512),
                                       (1 / 512, 1 / 1024, -1 / 2048),
                                       (1 / 1024, -1 / 512, 1 / 512)])
def test_basal_load_carries_the_centripetal_term(k1, k12, k2):
    # A layer moving over a curved bed presses on it with gravity's normal component plus
    # the centripetal force of following the bed: the normal acceleration of a path on the
    # surface is the second fundamental form of its velocity, II(V, V) = c (v . Hess(b) v)
    # for the bed z = b(x, y), c = 1/sqrt(1 + |grad b|^2) and horizontal velocity v.  The
    # scheme's curvature term (Flux15, L/K:2495-2527) is Ac = V . (dn/dt along the path)
    # = -II(V, V), so on the quadratic bed b = (k1 x^2 + 2 k12 x y + k2 y^2)/2 (exact in
    # float; its central differences are exact) with uniform velocities, Ac = -c v'Hv for
    # each phase, to the second-order error of the differentiated normal, and the basal
    # loads of Flux16 (L/K:2530-2559) follow with Det = 1/c:
    #   Npress2 = Det hs (c - Ac), Npress1 = Det hs ((1 - rr) c - Ac + rr AcF),
    #   NpressF = Det hf (c - AcF).
    n, dx = 48, 4.0
    i = np.arange(n) - n // 2
    X, Y = i[None, :] * dx, i[:, None] * dx
    topo = 100.0 + 0.5 * (k1 * X * X + 2 * k12 * X * Y + k2 * Y * Y)
    assert np.array_equal(f32(topo), topo)
    sc = S.Scenario(mode=0, topo=topo, depth=np.full((n, n), 2.0), cellsize=dx, phiS0=0.6,
                    delta0=20.0)
    o = Oracle(sc)
    u = o.get("u")
    us, vs, uf, vf = 30.0, -20.0, -25.0, 15.0               # ~30 m/s, scaled units
    u[1], u[2] = (u[0] + H0) * us, (u[0] + H0) * vs
    u[4], u[5] = (u[3] + H0) * uf, (u[3] + H0) * vf
    o.set_u(u)
    o.stage_terms()
    Hs = np.array([[k1, k12], [k12, k2]]) / 10.0            # curvature per scaled length
    Xs = (np.arange(o.nxd) - 3 - n // 2)[None, :] * dx * 10.0
    Ys = (np.arange(o.nyd) - 3 - n // 2)[:, None] * dx * 10.0
    bx, by = Hs[0, 0] * Xs + Hs[0, 1] * Ys, Hs[1, 0] * Xs + Hs[1, 1] * Ys
    c = 1.0 / np.sqrt(1.0 + bx * bx + by * by)

    def vHv(a, b):
        return Hs[0, 0] * a * a + 2 * Hs[0, 1] * a * b + Hs[1, 1] * b * b
    Ac, AcF = -c * vHv(us, vs), -c * vHv(uf, vf)
    inner = (slice(8, -8), slice(8, -8))
    assert np.abs(Ac[inner] / c[inner]).max() > 0.03        # the term matters
    np.testing.assert_allclose(o.get("Ac")[inner], Ac[inner], rtol=2e-4)
    np.testing.assert_allclose(o.get("AcF")[inner], AcF[inner], rtol=2e-4)
    rr, hs, hf = sc.rho_ratio, u[0], u[3]
    np.testing.assert_allclose(o.get("Npress2")[inner], (hs * (c - Ac) / c)[inner], rtol=1e-4)
    np.testing.assert_allclose(o.get("Npress1")[inner],
                               (hs * ((1 - rr) * c - Ac + rr * AcF) / c)[inner], rtol=1e-4)
    np.testing.assert_allclose(o.get("NpressF")[inner], (hf * (c - AcF) / c)[inner], rtol=1e-4)
