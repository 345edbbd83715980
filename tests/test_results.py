"""result2Pb post-processing (CPU): the plot script's derived fields and ESRI ASCII export,
on a hand-made result directory in the run path's formats."""
import math
import os

import numpy as np

from moses_2pdf_amd import results as R
from util import read_asc_np

ROW = "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n"
HEAD = 'VARIABLES = "H", "phi", "Us", "Uf", "Vs", "Vf"\n '


def make_dir(d, NX=4, NY=3, dx=20.0):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "Info.dat"), "w") as f:
        f.write('VARIABLES = "x-point", "y-point", "dx", "dy", "xllcorner", "yllcorner", "TotalStep"\n ')
        f.write("\t%d\t\t %d\t %10.2f\t %10.2f\t %10.4f\t %10.4f\t %d" % (NX, NY, dx, dx, 100.0, 200.0, 3))
    with open(os.path.join(d, "DEM.dat"), "w") as f:
        f.write('VARIABLES = "x", "y", "z", "c", "S1", "S2"\n ')
        for i in range(NX):
            for j in range(NY):
                f.write(ROW % (i * dx, j * dx, 10.0 * i + j, 0.5, 0.1, 0.2))
    with open(os.path.join(d, "Time.dat"), "w") as f:
        f.write("".join("%20.4f" % t for t in (0.0, 30.0, 60.0)))
    H = {}
    for k in (1, 2, 3):
        H[k] = np.zeros((NX, NY))
        with open(os.path.join(d, "%03d.dat" % k), "w") as f:
            f.write(HEAD)
            for i in range(NX):
                for j in range(NY):
                    h = 0.005 if (i + j + k) % 3 == 0 else 0.1 * (i + 1) * k
                    H[k][i, j] = float("%10.4f" % h)
                    f.write(ROW % (h, 0.6, 3.0, 4.0, 4.0, 3.0))
    return H


def test_reads_and_plot_fields(tmp_path):
    d = str(tmp_path / "r")
    H = make_dir(d)
    info = R.read_info(d)
    assert (info["NX"], info["NY"], info["dx"], info["outputs"]) == (4, 3, 20.0, 3)
    dem = R.read_dem(d, info)
    assert dem["x"][2, 1] == 40.0 and dem["y"][2, 1] == 20.0 and dem["z"][3, 2] == 32.0
    t = R.read_times(d)
    assert np.allclose(t, np.array([0.0, 30.0, 60.0]) * math.sqrt(0.1 / 9.8))
    res = R.read_result(d, 2, info)
    assert np.array_equal(res["H"], H[2])
    pf = R.plot_fields(res, dem)
    wet = H[2] >= 0.01
    assert np.isnan(pf["depth"][~wet]).all()
    assert np.allclose(pf["depth"][wet], H[2][wet] / 0.5)      # H / c
    assert np.allclose(pf["speedS"][wet], 5.0) and np.allclose(pf["speedF"][wet], 5.0)
    ext = R.flow_extent(d, info, [1, 2, 3])
    assert np.array_equal(ext, ((H[1] > 0.01) | (H[2] > 0.01) | (H[3] > 0.01)).astype(float))


def test_esri_export_reads_back_onto_the_grid(tmp_path):
    d = str(tmp_path / "r")
    H = make_dir(d)
    files = R.export_gis(d, str(tmp_path / "gis"))
    assert len(files) == 3 * 4 + 1
    p = str(tmp_path / "gis" / "depth_002.asc")
    hdr = open(p).read().split("\n")[:6]
    assert hdr[0].split() == ["ncols", "4"] and hdr[1].split() == ["nrows", "3"]
    assert float(hdr[2].split()[1]) == 100.0 and float(hdr[4].split()[1]) == 20.0
    g, cell = read_asc_np(p, False)           # the reader the input DEMs go through
    want = np.where(H[2] >= 0.01, H[2] / 0.5, R.NODATA).T    # line j = y index j
    assert g.shape == (3, 4) and np.allclose(g, want, atol=1e-4)


def test_figure(tmp_path):
    d = str(tmp_path / "r")
    make_dir(d)
    out = str(tmp_path / "fig.png")
    R.plot_png(d, 2, out)
    assert os.path.getsize(out) > 1000
