"""Post-processing of a result2Pb/ directory: the fields the reference's plot script draws,
and ESRI ASCII grid export.

Reads the files the run path writes (Info.dat, DEM.dat, Time.dat, NNN.dat; formats of
adNOC_2Pb_runKernels.cu:543-596, 1260-1324).  Records are x-major (the outer loop runs
over x), so a column of a record file reshapes to [NX, NY].

Plot fields (MoSES_2PDF_landslide/DataProcess/plotResult/plot_2PbGPU.py):
  - cells with H < 0.01 m are masked (NaN);
  - the plotted depth is H / c (depth normal to the terrain, c = cos of the slope);
  - solid and fluid speeds are |(Us, Vs)| and |(Uf, Vf)|;
  - the flow index marks cells where H > 0.01 at any plotted output;
  - output times are Time.dat times (scaled units) * sqrt(0.1 / 9.8) seconds.

GIS export: one ESRI ASCII grid per field and output.  The header is ncols NX, nrows NY,
xllcorner, yllcorner (from Info.dat), cellsize dx, NODATA_value -9999.  Data lines run
over y as the input DEM's lines do (file line j is grid row j), so a grid written here
reads back onto the DEM it came from.  The reference's own converter ships only as a
binary, so its exact output format is not reproduced (unpinned).

    python -m moses_2pdf_amd.results <result2Pb dir> [--gis OUT] [--png OUT] [--steps 1,5]
"""
from __future__ import annotations

import argparse
import math
import os
from typing import Dict, Iterable, List, Optional

import numpy as np

TIME_SCALE = math.sqrt(0.1 / 9.8)   # scaled time -> seconds (plot_2PbGPU.py)
H_MIN = 0.01                        # plotting mask, metres
NODATA = -9999.0


def read_info(d: str) -> Dict[str, float]:
    v = np.loadtxt(os.path.join(d, "Info.dat"), skiprows=1)
    return {"NX": int(v[0]), "NY": int(v[1]), "dx": float(v[2]), "dy": float(v[3]),
            "xll": float(v[4]), "yll": float(v[5]), "outputs": int(v[6])}


def _columns(path: str, n: int, ncol: int = 6) -> np.ndarray:
    a = np.loadtxt(path, skiprows=1)
    a = a.reshape(-1, ncol)
    if a.shape[0] != n:
        raise ValueError(f"{path}: {a.shape[0]} records, expected {n}")
    return a


def read_dem(d: str, info: Dict[str, float]) -> Dict[str, np.ndarray]:
    NX, NY = info["NX"], info["NY"]
    a = _columns(os.path.join(d, "DEM.dat"), NX * NY)
    return {k: a[:, q].reshape(NX, NY) for q, k in enumerate(("x", "y", "z", "c", "S1", "S2"))}


def read_times(d: str) -> np.ndarray:
    """Output times in seconds."""
    return np.atleast_1d(np.loadtxt(os.path.join(d, "Time.dat"))) * TIME_SCALE


def read_result(d: str, k: int, info: Dict[str, float]) -> Dict[str, np.ndarray]:
    NX, NY = info["NX"], info["NY"]
    a = _columns(os.path.join(d, "%03d.dat" % k), NX * NY)
    return {n: a[:, q].reshape(NX, NY) for q, n in enumerate(("H", "phi", "Us", "Uf", "Vs", "Vf"))}


def plot_fields(res: Dict[str, np.ndarray], dem: Dict[str, np.ndarray],
                h_min: float = H_MIN) -> Dict[str, np.ndarray]:
    """The masked fields of the plot script: depth H/c, phi, solid and fluid speed."""
    H = res["H"]
    mask = np.where(H < h_min, np.nan, 1.0)
    return {"depth": (H / dem["c"]) * mask, "phi": res["phi"] * mask,
            "speedS": np.sqrt(res["Us"] ** 2 + res["Vs"] ** 2) * mask,
            "speedF": np.sqrt(res["Uf"] ** 2 + res["Vf"] ** 2) * mask}


def flow_extent(d: str, info: Dict[str, float], steps: Iterable[int],
                h_min: float = H_MIN) -> np.ndarray:
    """1 where H > h_min at any of the given outputs, else 0 ([NX, NY])."""
    idx = np.zeros((info["NX"], info["NY"]))
    for k in steps:
        idx[read_result(d, k, info)["H"] > h_min] = 1
    return idx


def write_esri_ascii(path: str, field_xy: np.ndarray, info: Dict[str, float]) -> None:
    """Writes a [NX, NY] field as an ESRI ASCII grid (NaN -> NODATA_value)."""
    g = np.asarray(field_xy, np.float64).T          # [NY, NX]: line j holds y index j
    g = np.where(np.isfinite(g), g, NODATA)
    with open(path, "w") as f:
        f.write(f"ncols        {info['NX']}\n")
        f.write(f"nrows        {info['NY']}\n")
        f.write(f"xllcorner    {info['xll']:.6f}\n")
        f.write(f"yllcorner    {info['yll']:.6f}\n")
        f.write(f"cellsize     {info['dx']:.6f}\n")
        f.write(f"NODATA_value  {NODATA:.0f}\n")
        for j in range(g.shape[0]):
            f.write(" ".join("%.4f" % v for v in g[j]) + "\n")


def export_gis(d: str, out: str, steps: Optional[List[int]] = None,
               fields=("depth", "phi", "speedS", "speedF")) -> List[str]:
    """Grids of the plot fields for each output step (default: all written outputs)."""
    info = read_info(d)
    dem = read_dem(d, info)
    os.makedirs(out, exist_ok=True)
    if steps is None:
        steps = [k for k in range(1, info["outputs"] + 1) if os.path.exists(os.path.join(d, "%03d.dat" % k))]
    files = []
    for k in steps:
        pf = plot_fields(read_result(d, k, info), dem)
        for name in fields:
            p = os.path.join(out, f"{name}_{k:03d}.asc")
            write_esri_ascii(p, pf[name], info)
            files.append(p)
    p = os.path.join(out, "flow_extent.asc")
    write_esri_ascii(p, flow_extent(d, info, steps), info)
    files.append(p)
    return files


def plot_png(d: str, k: int, out_png: str, field: str = "depth") -> None:
    """One figure as the plot script draws it: terrain contours and the masked field."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    info = read_info(d)
    dem = read_dem(d, info)
    pf = plot_fields(read_result(d, k, info), dem)
    t = read_times(d)
    fig = plt.figure(figsize=(max(4, info["NX"] / 20), max(3, info["NY"] / 20)))
    ax = plt.gca()
    ax.set_aspect(1)
    plt.contour(dem["x"], dem["y"], dem["z"], 100, colors="black", linewidths=0.3)
    plt.pcolormesh(dem["x"], dem["y"], pf[field], shading="auto")
    plt.colorbar(shrink=0.82, aspect=20)
    if k - 1 < len(t):
        plt.text(0, dem["y"].max() * 1.02, "t = %.2f sec" % round(t[k - 1], 2))
    plt.xticks(())
    plt.yticks(())
    fig.savefig(out_png, dpi=80)
    plt.close(fig)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("result_dir")
    ap.add_argument("--gis", help="write ESRI ASCII grids into this directory")
    ap.add_argument("--png", help="write figures into this directory")
    ap.add_argument("--steps", help="comma-separated output numbers (default: all)")
    a = ap.parse_args(argv)
    steps = [int(s) for s in a.steps.split(",")] if a.steps else None
    if a.gis:
        for p in export_gis(a.result_dir, a.gis, steps):
            print(p)
    if a.png:
        info = read_info(a.result_dir)
        os.makedirs(a.png, exist_ok=True)
        for k in steps or range(1, info["outputs"] + 1):
            if os.path.exists(os.path.join(a.result_dir, "%03d.dat" % k)):
                p = os.path.join(a.png, "Fig_H%d.png" % (k - 1))
                plot_png(a.result_dir, k, p)
                print(p)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
