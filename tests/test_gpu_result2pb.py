"""The run path end to end: par_List and data files in, result2Pb/ out, byte for byte.

A scenario is written as a par_List tree, run through RunKernels (m2p_open + m2p_run) and
through the MoSES_2PDF_landslide / _inflow executables, and every file of result2Pb/ is
compared with text formatted here, independently, from the oracle's states at the same
output times (formats of adNOC_2Pb_runKernels.cu:543-596, 1260-1324)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from moses_2pdf_amd import M2P_INFLOW, M2P_LANDSLIDE, RunKernels
from moses_2pdf_amd import scenarios as S
from oracle.oracle import Oracle, f32, pad
from util import df004

pytestmark = pytest.mark.gpu

ROW = "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n"
HEAD = 'VARIABLES = "H", "phi", "Us", "Uf", "Vs", "Vf"\n '


def coord(c, n, d):
    if c < 3:
        return (0.0 + 0 * d) - (3 - c) * d
    if c >= n + 3:
        return (0.0 + (n - 1) * d) + (c - (n + 3 - 1)) * d
    return 0.0 + (c - 3) * d


def result_text(u, nx, ny):
    nxd, nyd = nx + 6, ny + 6
    P = nxd * nyd
    u = u.reshape(-1)
    out = [HEAD]
    for i in range(3, nxd - 3):
        for j in range(3, nyd - 3):
            q = j * nxd + i
            hs, hf = u[q], u[3 * P + q]
            if hs > 0.00001:
                vs, us, phi = u[2 * P + q] / hs, u[P + q] / hs, u[6 * P + q]
            else:
                vs = us = phi = 0.0
            if hf > 0.00001:
                vf, uf = u[5 * P + q] / hf, u[4 * P + q] / hf
            else:
                vf = uf = 0.0
            out.append(ROW % (0.1 * (hs + hf), phi, us, uf, vs, vf))
    return "".join(out)


def expected_files(sc):
    o = Oracle(sc, outputs=True)
    nx, ny = sc.nx, sc.ny
    nxd, nyd = nx + 6, ny + 6
    dx = float(np.float32(sc.cellsize)) * 10
    files = {}
    topo = pad(f32(sc.topo))
    c, sv = o.get("cvalue"), o.get("svec")
    dem = ['VARIABLES = "x", "y", "z", "c", "S1", "S2"\n ']
    for i in range(3, nxd - 3):
        for j in range(3, nyd - 3):
            dem.append(ROW % (coord(i, nx, dx) * 0.1, coord(j, ny, dx) * 0.1, topo[j, i] * 0.1,
                              c[j, i], sv[0, j, i], sv[1, j, i]))
    files["DEM.dat"] = "".join(dem)
    depth = pad(sc.depth) if sc.mode == 0 else np.zeros((nyd, nxd))
    init = [HEAD]
    for i in range(3, nxd - 3):
        for j in range(3, nyd - 3):
            init.append(ROW % (0.1 * depth[j, i], sc.phiS0, 0.0, 0.0, 0.0, 0.0))
    files["001.dat"] = "".join(init)
    ntf = int(sc.total_sim / sc.each_out) + 1
    files["Info.dat"] = ('VARIABLES = "x-point", "y-point", "dx", "dy", "xllcorner", "yllcorner", '
                         '"TotalStep"\n ' + "\t%d\t\t %d\t %10.2f\t %10.2f\t %10.4f\t %10.4f\t %d"
                         % (nx, ny, dx * 0.1, dx * 0.1, 0.0, 0.0, ntf))
    outtime = [0.0] * (ntf + 1)
    steps = 0
    while not o.scalar("nstop") and steps < 100000:
        o.step(1)
        steps += 1
        k = int(o.scalar("lastWritten"))
        if k:
            outtime[k - 1] = o.scalar("stepStartTime")
            files["Time.dat"] = "".join("%20.4f" % t for t in outtime[:k])
            files["%03d.dat" % k] = result_text(o.get("u"), nx, ny)
    return files, steps


def check_dir(d, files):
    got = sorted(os.listdir(d))
    assert got == sorted(files), (got, sorted(files))
    for name, text in files.items():
        with open(os.path.join(d, name)) as f:
            assert f.read() == text, name


def cases():
    sc = S.landslide_plane(n=40)
    sc.topo = f32(sc.topo)        # survives the %.9g text round trip exactly
    sc.total_sim, sc.each_out = 6.0, 2.0
    sc2 = df004()
    sc2.total_sim, sc2.each_out = 20.0, 10.0
    return [("landslide", sc, M2P_LANDSLIDE), ("inflow", sc2, M2P_INFLOW)]


@pytest.mark.parametrize("name,sc,mode", cases(), ids=["landslide", "inflow"])
def test_run_writes_result2pb_byte_exact(tmp_path, name, sc, mode):
    root = str(tmp_path / name)
    S.write_case(sc, root)
    files, steps = expected_files(sc)
    assert steps > 3 and len([f for f in files if f[:3].isdigit()]) >= 3
    cwd = os.getcwd()
    os.chdir(root)
    try:
        RunKernels("./par_List", mode=mode).run("./result2Pb/")
    finally:
        os.chdir(cwd)
    check_dir(os.path.join(root, "result2Pb"), files)
    # post-processing reads the real output: GIS grids of every written output
    from moses_2pdf_amd import results as R
    grids = R.export_gis(os.path.join(root, "result2Pb"), os.path.join(root, "gis"))
    assert len(grids) == 4 * len([f for f in files if f[:3].isdigit()]) + 1
    # the drop-in executable, in the same working-directory contract
    exe = os.path.join(ROOT, "moses_2pdf_amd", "bin",
                       "MoSES_2PDF_landslide" if mode == M2P_LANDSLIDE else "MoSES_2PDF_inflow")
    if not os.path.exists(exe):
        pytest.skip("executables not built")
    for f in os.listdir(os.path.join(root, "result2Pb")):
        os.remove(os.path.join(root, "result2Pb", f))
    r = subprocess.run([exe], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Time used on GPU:" in r.stdout
    check_dir(os.path.join(root, "result2Pb"), files)
    # every tile computed, as the reference does (overlay dry_skip=0): the same files
    for f in os.listdir(os.path.join(root, "result2Pb")):
        os.remove(os.path.join(root, "result2Pb", f))
    r = subprocess.run([exe, "dry_skip=0", "./result2Pb/"], cwd=root, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    check_dir(os.path.join(root, "result2Pb"), files)


def test_missing_par_list_is_an_error_not_an_exit(tmp_path):
    from moses_2pdf_amd import M2PError
    with pytest.raises(M2PError):
        RunKernels(str(tmp_path / "nope"), mode=M2P_LANDSLIDE)
