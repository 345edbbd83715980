"""Seeded synthetic hazard scenarios in the reference's input formats.

Every generator returns plain numpy arrays in file orientation (``a[j, i]``: line ``j``
of an ESRI ASCII grid, value ``i`` on that line), in metres, plus the run parameters.
``write_*`` helpers emit the same scenario as a ``par_List`` tree, so the file reader,
the CPU oracle and the GPU path all see identical bytes.

Parameters follow the reference's par_List schema (MoSES_2PDF_landslide/par_List:1-9,
MoSES_2PDF_inflow/par_List:1-11): total time, output interval, delta0 [deg], Cd, N_R,
varTheta, phiS0, then file paths.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

# The reference's own material constants (par_List:3-6 of the landslide mode).
REF_DELTA0 = 16.0
REF_CD = 6.0
REF_NR = 268.0
REF_VARTHETA = 5.0


@dataclasses.dataclass
class Scenario:
    """A landslide or inflow scenario, in the reference's (unscaled, metres) units."""

    mode: int                      # 0 landslide, 1 inflow
    topo: np.ndarray               # [NY, NX] metres
    cellsize: float                # metres
    delta0: float = REF_DELTA0
    Cd: float = REF_CD
    N_R: float = REF_NR
    varTheta: float = REF_VARTHETA
    phiS0: float = 0.5
    total_sim: float = 0.0         # seconds, 0: fixed-step runs only
    each_out: float = 1.0
    depth: Optional[np.ndarray] = None   # [NY, NX] metres (landslide)
    phi: Optional[np.ndarray] = None     # [NY, NX] per-cell solid fraction (extension)
    # inflow mode: hydrograph samples and locations (I/R:297-331)
    inflow_t: Optional[np.ndarray] = None   # [nsamp] seconds
    inflow_h: Optional[np.ndarray] = None   # [nsamp, nloc] metres
    inflow_u: Optional[np.ndarray] = None   # [nsamp, nloc]
    inflow_v: Optional[np.ndarray] = None   # [nsamp, nloc]
    inflow_loc: Optional[np.ndarray] = None  # [nloc, 3] (NO, X, Y)
    # knobs the reference hard-codes (kernels.cuh:16, kernels.cu:1710, 2014-2017)
    cfl: float = 0.1
    rho_ratio: float = 1.420 / 2.60
    dt_min: Optional[float] = None
    dt_max: float = 2.0

    @property
    def nx(self) -> int:
        return int(self.topo.shape[1])

    @property
    def ny(self) -> int:
        return int(self.topo.shape[0])

    def dt_floor(self) -> float:
        if self.dt_min is not None:
            return self.dt_min
        return 0.001 if self.mode == 0 else 0.002


# ---------------------------------------------------------------------------------------
# terrain
# ---------------------------------------------------------------------------------------

def inclined_plane(nx: int, ny: int, dx: float, angle_deg: float = 35.0,
                   x_break: float = 3000.0, blend: float = 250.0) -> np.ndarray:
    """A plane of ``angle_deg`` falling in +x that blends linearly in slope (tan) over
    ``[x_break-blend, x_break+blend]`` into a flat runout at height 0."""
    x = np.arange(nx, dtype=np.float64) * dx
    t = math.tan(math.radians(angle_deg))
    a, b = x_break - blend, x_break + blend
    z = np.where(x >= b, 0.0,
                 np.where(x >= a, t * (b - x) ** 2 / (2 * (b - a)),
                          t * (a - x) + t * (b - a) / 2))
    return np.repeat(z[None, :], ny, axis=0)


def gaussian_bump(nx: int, ny: int, dx: float, x0: float, y0: float,
                  amp: float, sigma: float) -> np.ndarray:
    x = np.arange(nx, dtype=np.float64) * dx
    y = np.arange(ny, dtype=np.float64) * dx
    r2 = (x[None, :] - x0) ** 2 + (y[:, None] - y0) ** 2
    return amp * np.exp(-r2 / (2 * sigma * sigma))


def diamond_square(n_pow2: int, hurst: float, rng: np.random.Generator) -> np.ndarray:
    """Fractal surface on a (2^k+1)^2 lattice, displacement sigma_k = 2^(-H k)."""
    size = n_pow2 + 1
    z = np.zeros((size, size), dtype=np.float64)
    z[0, 0], z[0, -1], z[-1, 0], z[-1, -1] = rng.standard_normal(4)
    step = n_pow2
    scale = 1.0
    while step > 1:
        half = step // 2
        # diamond step
        c = (z[0:-1:step, 0:-1:step] + z[step::step, 0:-1:step] +
             z[0:-1:step, step::step] + z[step::step, step::step]) * 0.25
        z[half::step, half::step] = c + scale * rng.standard_normal(c.shape)
        # square step
        zp = np.pad(z, half, mode="edge")
        for (oy, ox) in ((0, half), (half, 0)):
            ys = np.arange(oy, size, step)
            xs = np.arange(ox, size, step)
            Y, X = np.meshgrid(ys + half, xs + half, indexing="ij")
            avg = (zp[Y - half, X] + zp[Y + half, X] + zp[Y, X - half] + zp[Y, X + half]) * 0.25
            z[np.ix_(ys, xs)] = avg + scale * rng.standard_normal(avg.shape)
        step = half
        scale *= 2.0 ** (-hurst)
    return z


def box3(z: np.ndarray) -> np.ndarray:
    zp = np.pad(z, 1, mode="edge")
    acc = np.zeros_like(z)
    for dy in range(3):
        for dx_ in range(3):
            acc += zp[dy:dy + z.shape[0], dx_:dx_ + z.shape[1]]
    return acc / 9.0


def fractal_dem(nx: int, ny: int, relief: float, hurst: float, seed: int,
                smooth: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = 1
    while n < max(nx, ny):
        n *= 2
    z = diamond_square(n, hurst, rng)[:ny, :nx]
    z = (z - z.min()) / max(z.max() - z.min(), 1e-300) * relief
    return box3(z) if smooth else z


def paraboloid(nx: int, ny: int, dx: float, xc: float, yc: float, volume: float,
               radius: float) -> np.ndarray:
    """h = h0 (1 - r^2/R^2), h0 = 2V/(pi R^2)."""
    x = np.arange(nx, dtype=np.float64) * dx
    y = np.arange(ny, dtype=np.float64) * dx
    r2 = (x[None, :] - xc) ** 2 + (y[:, None] - yc) ** 2
    h0 = 2.0 * volume / (math.pi * radius * radius)
    return np.maximum(0.0, h0 * (1.0 - r2 / (radius * radius)))


# ---------------------------------------------------------------------------------------
# the benchmark configurations
# ---------------------------------------------------------------------------------------

def landslide_plane(n: int = 256, seed: int = 1, dx: float = 20.0, phi: float = 0.6,
                    volume: float = 1.0e6, radius: float = 150.0,
                    release_xy=(1000.0, 2560.0)) -> Scenario:
    """Landslide on a 35 degree plane blending into a flat runout over 500 m, with a seeded
    Gaussian bump (30 m, sigma 200 m) and a paraboloid release of ``volume`` m^3.
    Grid ``n x n`` at ``dx`` metres (the 256^2 configuration at 20 m by default)."""
    rng = np.random.default_rng(seed)
    L = n * dx
    s = L / 5120.0  # positions scale with the domain
    bx = rng.uniform(1500.0, 2500.0) * s
    by = rng.uniform(1500.0, 3600.0) * s
    topo = inclined_plane(n, n, dx, x_break=3000.0 * s, blend=250.0 * s)
    topo = topo + gaussian_bump(n, n, dx, bx, by, 30.0, 200.0 * s)
    depth = paraboloid(n, n, dx, release_xy[0] * s, release_xy[1] * s, volume * s * s, radius * s)
    return Scenario(mode=0, topo=topo, depth=depth, cellsize=dx, phiS0=phi,
                    delta0=20.0, Cd=REF_CD, N_R=REF_NR, varTheta=REF_VARTHETA)


def landslide_fractal(n: int = 2048, seed: int = 2, dx: float = 20.0,
                      relief: float = 1500.0) -> Scenario:
    """Fractal DEM (diamond-square, Hurst 0.8, ``relief`` metres, 3-cell box filter) with
    three seeded paraboloid releases of 1e5..5e6 m^3 and per-release solid fraction in
    [0.4, 0.7] (the per-cell phi extension)."""
    rng = np.random.default_rng(seed)
    topo = fractal_dem(n, n, relief, 0.8, seed)
    depth = np.zeros_like(topo)
    phi = np.full_like(topo, 0.5)
    hi = np.argwhere(topo > np.median(topo))
    for _ in range(3):
        vol = math.exp(rng.uniform(math.log(1e5), math.log(5e6)))
        h0 = 20.0
        radius = math.sqrt(2.0 * vol / (math.pi * h0))
        j, i = hi[rng.integers(len(hi))]
        p = paraboloid(n, n, dx, i * dx, j * dx, vol, radius)
        ph = rng.uniform(0.4, 0.7)
        mask = p > depth
        depth = np.where(mask, p, depth)
        phi = np.where(mask & (p > 0), ph, phi)
    return Scenario(mode=0, topo=topo, depth=depth, phi=phi, cellsize=dx, phiS0=0.5,
                    delta0=20.0)


def self_affine_relief(n: int, dx: float = 20.0, relief_2048: float = 1500.0,
                       hurst: float = 0.8) -> float:
    """Relief of an n-cell square that has the slope statistics of a 2048-cell one with
    ``relief_2048`` metres: relief scales as extent^H for a self-affine surface."""
    return relief_2048 * (n * dx / (2048 * 20.0)) ** hurst


@dataclasses.dataclass
class MemberParams:
    """One ensemble member: grid side and the sampled physical parameters."""
    n: int
    seed: int
    delta0: float          # basal friction angle [deg]
    internal: float        # internal friction angle [deg] (sampled; this model has none)
    volume: float          # release volume [m^3]
    phi: float             # solid volume fraction of the release


def latin_hypercube(count: int, seed: int, ranges) -> np.ndarray:
    """``count`` samples of len(ranges) variables, one per stratum of each, strata paired
    by independent random permutations; each range is (lo, hi, 'lin' | 'log')."""
    rng = np.random.default_rng(seed)
    out = np.empty((count, len(ranges)))
    for k, (lo, hi, kind) in enumerate(ranges):
        u = (rng.permutation(count) + rng.uniform(size=count)) / count
        if kind == "log":
            out[:, k] = np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
        else:
            out[:, k] = lo + u * (hi - lo)
    return out


def ensemble_members(count: int = 256, seed: int = 4, sizes=(128, 256, 512, 1024),
                     weights=(4, 3, 2, 1)):
    """The batched-ensemble population: sides drawn from ``sizes`` with ``weights``, basal
    friction 15-30 deg, internal friction 30-40 deg, release volume 1e5-5e6 m^3 (log),
    solid fraction 0.4-0.7, by Latin hypercube."""
    rng = np.random.default_rng(seed)
    w = np.asarray(weights, np.float64)
    ns = rng.choice(np.asarray(sizes), size=count, p=w / w.sum())
    lhs = latin_hypercube(count, seed + 1, [(15.0, 30.0, "lin"), (30.0, 40.0, "lin"),
                                            (1e5, 5e6, "log"), (0.4, 0.7, "lin")])
    return [MemberParams(int(ns[k]), seed * 100003 + k, *map(float, lhs[k])) for k in range(count)]


def ensemble_scenario(m: MemberParams, dx: float = 20.0) -> Scenario:
    """Landslide of one ensemble member: seeded fractal DEM with self-affine relief
    (the slope statistics of the 2048-cell, 1500 m DEM), one paraboloid release of
    ``m.volume`` at a seeded high point (20 m peak depth, radius capped at a quarter of
    the domain), solid fraction ``m.phi``."""
    rng = np.random.default_rng(m.seed)
    n = m.n
    topo = fractal_dem(n, n, self_affine_relief(n, dx), 0.8, m.seed)
    radius = min(math.sqrt(2.0 * m.volume / (math.pi * 20.0)), 0.25 * n * dx)
    rc = int(math.ceil(radius / dx)) + 2
    inner = topo[rc:n - rc, rc:n - rc]
    cand = np.argwhere(inner >= np.quantile(inner, 0.75)) + rc
    j, i = cand[rng.integers(len(cand))]
    depth = paraboloid(n, n, dx, i * dx, j * dx, m.volume, radius)
    return Scenario(mode=0, topo=topo, depth=depth, cellsize=dx, phiS0=m.phi, delta0=m.delta0)


def inflow_valley(nx: int = 1024, ny: int = 4096, seed: int = 3, dx: float = 20.0,
                  nloc: int = 16, t_end: float = 1200.0) -> Scenario:
    """V-shaped valley along +y: thalweg slope 25 deg at row 0 falling linearly to 3 deg,
    side slopes 30 deg, fractal roughness of 5 m amplitude.  Triangular hydrograph
    (peak 2000 m^3/s at 300 s, ending at 900 s) on ``nloc`` upstream-edge cells (Y = 0,
    so each injects a 1x6 strip)."""
    rng = np.random.default_rng(seed)
    y = np.arange(ny, dtype=np.float64) * dx
    ang = np.radians(25.0 + (3.0 - 25.0) * np.arange(ny) / max(ny - 1, 1))
    slope = np.tan(ang)
    zt = np.concatenate([[0.0], np.cumsum(0.5 * (slope[1:] + slope[:-1]) * dx)])
    zt = zt[-1] - zt  # falls along +y
    xc = (nx // 2) * dx
    x = np.arange(nx, dtype=np.float64) * dx
    topo = zt[:, None] + math.tan(math.radians(30.0)) * np.abs(x[None, :] - xc)
    rough = fractal_dem(nx, ny, 1.0, 0.8, seed, smooth=False)
    rough = (rough - rough.mean())
    rough = rough / max(np.abs(rough).max(), 1e-300) * 5.0
    topo = topo + rough
    t = np.arange(0.0, t_end + 1e-9, 10.0)
    Q = np.where(t <= 300.0, 2000.0 * t / 300.0, np.where(t <= 900.0, 2000.0 * (900.0 - t) / 600.0, 0.0))
    vel = 5.0
    h = Q / (nloc * dx * vel)
    ic = nx // 2 + 3  # padded column index of the thalweg
    X = ic - nloc // 2 + np.arange(nloc)
    loc = np.stack([np.arange(1, nloc + 1), X, np.zeros(nloc)], axis=1).astype(np.float64)
    H = np.repeat(h[:, None], nloc, axis=1)
    U = np.zeros_like(H)
    V = np.full_like(H, vel)
    del rng
    return Scenario(mode=1, topo=topo, cellsize=dx, phiS0=0.5, delta0=20.0,
                    inflow_t=t, inflow_h=H, inflow_u=U, inflow_v=V, inflow_loc=loc)


# ---------------------------------------------------------------------------------------
# writers (the reference's on-disk formats)
# ---------------------------------------------------------------------------------------

def write_asc(path: str, a: np.ndarray, cellsize: float, nodata: bool = True,
              fmt: str = "%.17g") -> None:
    ny, nx = a.shape
    with open(path, "w") as f:
        f.write(f"ncols        {nx}\n")
        f.write(f"nrows        {ny}\n")
        f.write("xllcorner    0.000000000000\n")
        f.write("yllcorner    0.000000000000\n")
        f.write(f"cellsize     {cellsize:.12f}\n")
        if nodata:
            f.write("NODATA_value  -3.4028234663852885981e+38\n")
        for j in range(ny):
            f.write(" " + " ".join(fmt % v for v in a[j]) + "\n")


def write_case(sc: Scenario, root: str) -> str:
    """Writes ``root/par_List`` and its data files; returns the par_List path."""
    os.makedirs(os.path.join(root, "Data"), exist_ok=True)
    os.makedirs(os.path.join(root, "result2Pb"), exist_ok=True)
    write_asc(os.path.join(root, "Data", "topo.asc"), sc.topo, sc.cellsize, fmt="%.9g")
    lines = [f"{sc.total_sim!r}      # Total simulation time (sec)",
             f"{sc.each_out!r}       # Each output time (sec)",
             f"{sc.delta0!r} # delta0 (friction angle of solid constituent) in degree",
             f"{sc.Cd!r}    # value of Cd",
             f"{sc.N_R!r}   # value of N_R",
             f"{sc.varTheta!r}    # value of varTheta",
             f"{sc.phiS0!r}    # initial value of solid volume fraction",
             "./Data/topo.asc\t\t# topography data (DEM)"]
    if sc.mode == 0:
        write_asc(os.path.join(root, "Data", "init.asc"), sc.depth, sc.cellsize)
        lines.append("./Data/init.asc\t\t# initial depth")
    else:
        names = ["h", "u", "v"]
        arrs = [sc.inflow_h, sc.inflow_u, sc.inflow_v]
        nloc = sc.inflow_loc.shape[0]
        for nm, arr in zip(names, arrs):
            with open(os.path.join(root, "Data", f"in_{nm}_t.dat"), "w") as f:
                f.write("time(sec)\t" + "\t".join(f"location{i + 1}" for i in range(nloc)) + "\n")
                for k, t in enumerate(sc.inflow_t):
                    f.write("%.9g\t" % t + "\t".join("%.9g" % x for x in arr[k]) + "\n")
        with open(os.path.join(root, "Data", "in_location.dat"), "w") as f:
            f.write("NO\t\tX\t\tY\n")
            for r in sc.inflow_loc:
                f.write("%f\t%f\t%f\n" % tuple(r))
        lines += ["./Data/in_h_t.dat\t# initial depth (H)",
                  "./Data/in_u_t.dat\t# initial velocity (u)",
                  "./Data/in_v_t.dat\t# initial velocity (v)",
                  "./Data/in_location.dat\t# inflow point location"]
    p = os.path.join(root, "par_List")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n")
    return p
