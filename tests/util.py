"""Shared helpers for the tests: scenarios from the reference's data files and
comparison of two state arrays."""
from __future__ import annotations

import os

import numpy as np

from moses_2pdf_amd import scenarios as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REFDATA = os.path.join(GOLDEN, "reference_data")


def read_asc_np(path, as_float):
    """Independent reader of an ESRI ASCII grid with the reference's rules
    (adNOC_2Pb_runKernels.cu:137-278)."""
    with open(path) as f:
        lines = f.read().split("\n")
    hdr = [l.split() for l in lines[:6]]
    nx, ny = int(hdr[0][1]), int(hdr[1][1])
    cell = float(np.float32(hdr[4][1]))
    start = 6 if hdr[5][0] == "NODATA_value" else 5
    rows = [lines[start + j].split()[:nx] for j in range(ny)]
    a = np.array(rows, dtype=np.float64)
    if as_float:
        a = np.array(rows, dtype=np.float32).astype(np.float64)
    return a, cell


def hsiaolin(delta0=16.0, phi=0.5):
    topo, cell = read_asc_np(os.path.join(REFDATA, "shiaolin_topo_20m.asc"), True)
    depth, _ = read_asc_np(os.path.join(REFDATA, "shiaolin_init20m.asc"), False)
    return S.Scenario(mode=0, topo=topo, depth=depth, cellsize=cell, delta0=delta0, phiS0=phi,
                      total_sim=180.0, each_out=3.0)


def df004(phi=0.5):
    topo, cell = read_asc_np(os.path.join(REFDATA, "00_DebrisFlow_dem_20m.asc"), True)

    def tab(name):
        with open(os.path.join(REFDATA, name)) as f:
            rows = [l.split("\t") for l in f.read().split("\n")[1:] if l.strip()]
        return np.array(rows, dtype=np.float32).astype(np.float64)

    h, u, v = tab("00_h_t.dat"), tab("00_u_t.dat"), tab("00_v_t.dat")
    loc = tab("00_location.dat")
    return S.Scenario(mode=1, topo=topo, cellsize=cell, delta0=16.0, phiS0=phi,
                      total_sim=720.0, each_out=10.0, inflow_t=h[:, 0], inflow_h=h[:, 1:],
                      inflow_u=u[:, 1:], inflow_v=v[:, 1:], inflow_loc=loc)


def interior(u):
    return u[..., 3:-3, 3:-3]


def compare(a, b, name=""):
    """Max relative difference (|a-b| / max(|a|,|b|, tiny)) and the count of cells whose
    values differ at all (+0.0 and -0.0 compare equal)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    d = np.abs(a - b)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-300)
    rel = np.where(d == 0, 0.0, d / scale)
    return float(rel.max()) if rel.size else 0.0, int(np.count_nonzero(a != b))
