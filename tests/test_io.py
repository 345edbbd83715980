"""The reference's file formats through the product's readers and writers (CPU only).

tests/io_check.cpp drives moses_2pdf_amd/csrc/m2p_io.cpp (no GPU).  Expected values come
from an independent parse (numpy, Python's float formatting) of the reference's own data
files, copied under tests/golden/reference_data/.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from util import REFDATA, read_asc_np


@pytest.fixture(scope="module")
def io_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("io") / "io_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "io_check.cpp"),
                    os.path.join(ROOT, "moses_2pdf_amd", "csrc", "m2p_io.cpp")], check=True)
    return exe


def run(exe, *args, cwd=None):
    r = subprocess.run([exe, *args], capture_output=True, text=True, cwd=cwd)
    return r.returncode, r.stdout


def test_par_list_landslide(io_check):
    rc, out = run(io_check, "par", "0", os.path.join(REFDATA, "par_List.landslide"))
    assert rc == 0, out
    lines = out.split("\n")
    assert [float(x) for x in lines[0].split()] == [180, 3, 16.0, 6.0, 268.0, 5.0, 0.5]
    assert lines[1] == "./Data/hsiaolin_landslide_asc/shiaolin_topo_20m.asc"
    assert lines[2] == "./Data/hsiaolin_landslide_asc/shiaolin_init20m.asc"


def test_par_list_inflow(io_check):
    rc, out = run(io_check, "par", "1", os.path.join(REFDATA, "par_List.inflow"))
    assert rc == 0, out
    lines = out.split("\n")
    assert [float(x) for x in lines[0].split()] == [720, 10, 16.0, 6.0, 268.0, 5.0, 0.5]
    assert lines[1].endswith("00_DebrisFlow_dem_20m.asc")
    assert lines[3].endswith("00_h_t.dat") and lines[6].endswith("00_location.dat")


def test_par_list_errors(io_check, tmp_path):
    rc, out = run(io_check, "par", "0", str(tmp_path / "missing"))
    assert rc == 1 and "missing" in out
    p = tmp_path / "short"
    p.write_text("180 # total\n3\n")
    rc, out = run(io_check, "par", "0", str(p))
    assert rc == 1 and "line 3" in out


@pytest.mark.parametrize("name,as_float,nodata", [("shiaolin_topo_20m.asc", 1, True),
                                                   ("shiaolin_init20m.asc", 0, True),
                                                   ("00_DebrisFlow_dem_20m.asc", 1, False)])
def test_asc_reader(io_check, name, as_float, nodata):
    path = os.path.join(REFDATA, name)
    rc, out = run(io_check, "asc", path, str(as_float))
    assert rc == 0, out[:200]
    lines = out.split("\n")
    nx, ny, xll, yll, cell = lines[0].split()
    a, c = read_asc_np(path, bool(as_float))
    assert (int(nx), int(ny)) == (a.shape[1], a.shape[0])
    hdr = open(path).read().split("\n")
    assert float(xll) == float(np.float32(hdr[2].split()[1]))   # stof
    assert float(cell) == c
    vals = np.array([float(x) for x in lines[1:1 + a.size]])
    assert np.array_equal(vals, a.ravel())
    assert (hdr[5].split()[0] == "NODATA_value") == nodata


def test_hydrograph_reader(io_check, tmp_path):
    for f in os.listdir(REFDATA):
        shutil.copy(os.path.join(REFDATA, f), tmp_path / f)
    par = tmp_path / "par_List"
    par.write_text(open(os.path.join(REFDATA, "par_List.inflow")).read()
                   .replace("./Data/DF004_3hr_6min/", "./"))
    rc, out = run(io_check, "hydro", "par_List", cwd=str(tmp_path))
    assert rc == 0, out[:300]
    v = out.split("\n")
    nsamp, nloc = map(int, v[0].split())
    assert (nsamp, nloc) == (71, 8)
    nums = np.array([float(x) for x in v[1:] if x])

    def tab(n):
        rows = [l.split("\t") for l in open(os.path.join(REFDATA, n)).read().split("\n")[1:] if l.strip()]
        return np.array(rows, dtype=np.float32).astype(np.float64)
    h, u, vv, loc = tab("00_h_t.dat"), tab("00_u_t.dat"), tab("00_v_t.dat"), tab("00_location.dat")
    k = 0
    assert np.array_equal(nums[k:k + nsamp], h[:, 0]); k += nsamp
    for t in (h, u, vv):
        assert np.array_equal(nums[k:k + nsamp * nloc], t[:, 1:].ravel()); k += nsamp * nloc
    assert np.array_equal(nums[k:k + 3 * nloc], loc.ravel())
    assert h[0, 0] == 1.0 and h[-1, 0] == 700.0


def expected_result(u, nx, ny):
    """Reference NNN.dat text (adNOC_2Pb_runKernels.cu:1287-1313), x-major."""
    nxd, nyd = nx + 6, ny + 6
    P = nxd * nyd
    out = ['VARIABLES = "H", "phi", "Us", "Uf", "Vs", "Vf"\n ']
    for i in range(3, nxd - 3):
        for j in range(3, nyd - 3):
            q = j * nxd + i
            hs, hf = u[q], u[3 * P + q]
            H = 0.1 * (hs + hf)
            if hs > 0.00001:
                vs, us, phi = u[2 * P + q] / hs, u[P + q] / hs, u[6 * P + q]
            else:
                vs = us = phi = 0.0
            if hf > 0.00001:
                vf, uf = u[5 * P + q] / hf, u[4 * P + q] / hf
            else:
                vf = uf = 0.0
            out.append("%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n" % (H, phi, us, uf, vs, vf))
    return "".join(out)


def test_result_writers_byte_exact(io_check, tmp_path):
    rc, out = run(io_check, "write", str(tmp_path))
    assert rc == 0, out
    nx, ny = 4, 3
    P = (nx + 6) * (ny + 6)
    u = np.sin(0.37 * np.arange(7 * P, dtype=np.float64)) * 3.0
    idx = np.arange(P)
    u[:P][idx % 5 == 0] = 0.0
    u[3 * P:4 * P][idx % 7 == 0] = 1e-6
    assert (tmp_path / "007.dat").read_text() == expected_result(u, nx, ny)
    # %03d numbering (L/R:1273) keeps counting past 999
    assert (tmp_path / "1234.dat").read_text() == expected_result(u, nx, ny)
    assert (tmp_path / "Time.dat").read_text() == "%20.4f%20.4f%20.4f" % (0.0, 30.25, 60.5)
    assert (tmp_path / "Info.dat").read_text() == (
        'VARIABLES = "x-point", "y-point", "dx", "dy", "xllcorner", "yllcorner", "TotalStep"\n '
        "\t%d\t\t %d\t %10.2f\t %10.2f\t %10.4f\t %10.4f\t %d" % (185, 111, 20.0, 20.0, 212862.0625,
                                                                  2561457.5, 61))
    init = (tmp_path / "001.dat").read_text().split("\n")
    assert init[0] == 'VARIABLES = "H", "phi", "Us", "Uf", "Vs", "Vf"'
    # x-major: the second line is cell (col 3, row 3), the third (col 3, row 4)
    nxd = nx + 6
    assert init[1] == " " + "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f" % (
        0.1 * 0.25 * (3 * nxd + 3), 0.55, 0, 0, 0, 0)
    assert init[2] == "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f" % (
        0.1 * 0.25 * (4 * nxd + 3), 0.55, 0, 0, 0, 0)
