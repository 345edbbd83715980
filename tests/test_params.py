"""Validation of the run parameters and of the overlay string (CPU: every call here is
rejected before any device work).  The reference reads its par_List with sscanf and checks
nothing (MoSES_2PDF_landslide/adNOC_2Pb_runKernels.cu:11-48); the library returns
M2P_EINVAL with the reason instead of running a case it cannot run."""
import ctypes
import os

import numpy as np
import pytest

from moses_2pdf_amd import engine as E
from moses_2pdf_amd import scenarios as S

EINVAL = -1


def _open_grid(mode=0, **over):
    L = E.lib()
    p = E.default_params(mode)
    for k, v in over.items():
        setattr(p, k, v)
    n = 8
    topo = np.zeros(n * n)
    depth = np.ones(n * n)
    h = ctypes.c_void_p()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if mode == 0:
        rc = L.m2p_open_grid(0, n, n, 20.0, dp(topo), dp(depth), None, ctypes.byref(p), 0, 0,
                             None, None, None, None, None, ctypes.byref(h))
    else:
        t = np.array([1.0, 2.0])
        hh = np.ones(2)
        loc = np.array([1.0, 3.0, 0.0])
        rc = L.m2p_open_grid(1, n, n, 20.0, dp(topo), None, None, ctypes.byref(p), 2, 1,
                             dp(t), dp(hh), dp(hh), dp(hh), dp(loc), ctypes.byref(h))
    msg = L.m2p_last_error(h).decode() if h.value else ""
    if h.value:
        L.m2p_close(h)
    return rc, msg


@pytest.mark.parametrize("field,value,word", [
    ("cfl", 0.0, "cfl"), ("cfl", -0.1, "cfl"), ("cfl", 1.5, "cfl"), ("cfl", float("nan"), "cfl"),
    ("rho_ratio", 1.0, "rho_ratio"), ("rho_ratio", -0.2, "rho_ratio"),
    ("dt_min", 0.0, "dt_min"), ("dt_min", -1.0, "dt_min"),
    ("dt_max", 0.0005, "dt_max"),            # below the landslide dt_min 0.001
    ("dt_max", float("inf"), "dt_max"),
    ("N_R", 0.0, "N_R"), ("Cd", -1.0, "Cd"), ("varTheta", -1.0, "varTheta"),
    ("delta0", 90.0, "delta0"), ("delta0", -5.0, "delta0"),
    ("phiS0", 1.5, "phiS0"), ("phiS0", -0.1, "phiS0"),
    ("each_out", 0.0, "each_out"), ("total_sim", -1.0, "total_sim"),
    ("precision", 7, "precision"), ("dry_skip", 2, "dry_skip"),
    ("metric_form", 2, "metric_form"), ("metric_form", -1, "metric_form"),
])
def test_out_of_range_parameters_are_rejected(field, value, word):
    rc, msg = _open_grid(0, **{field: value})
    assert rc == EINVAL
    assert word in msg and "invalid parameters" in msg


def test_inflow_needs_a_positive_solid_fraction():
    rc, msg = _open_grid(1, phiS0=0.0)
    assert rc == EINVAL and "phiS0" in msg and "inflow" in msg


def _case(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)    # par_List paths resolve against the current directory
    sc = S.Scenario(mode=0, topo=np.zeros((8, 8)), depth=np.ones((8, 8)), cellsize=20.0,
                    total_sim=10.0, each_out=1.0)
    return S.write_case(sc, str(tmp_path))


@pytest.mark.parametrize("overlay", ["cfl=abc", "cfl=0.4x", "cfl=", "dt_min=nan",
                                     "dt_max=inf", "rho_ratio=1e", "unknown=1", "cfl",
                                     "precision=fp16", "dry_skip=yes", "metric_form=2",
                                     "metric_form=stored"])
def test_malformed_overlay_is_rejected(tmp_path, monkeypatch, overlay):
    par = _case(tmp_path, monkeypatch)
    with pytest.raises(E.M2PError) as e:
        E.Simulation(par_list=par, mode=0, overlay=overlay)
    assert "(-1)" in str(e.value) and "overlay" in str(e.value)


@pytest.mark.parametrize("overlay,word", [("cfl=2", "cfl"), ("dt_min=0.5;dt_max=0.1", "dt_max"),
                                          ("rho_ratio=1.2", "rho_ratio")])
def test_out_of_range_overlay_is_rejected(tmp_path, monkeypatch, overlay, word):
    par = _case(tmp_path, monkeypatch)
    with pytest.raises(E.M2PError) as e:
        E.Simulation(par_list=par, mode=0, overlay=overlay)
    assert "(-1)" in str(e.value) and word in str(e.value)


def test_malformed_par_list_value_is_rejected(tmp_path, monkeypatch):
    par = _case(tmp_path, monkeypatch)
    lines = open(par).read().split("\n")
    lines[4] = "0.0   # value of N_R"           # N_R = 0 divides the viscosity
    open(par, "w").write("\n".join(lines))
    with pytest.raises(E.M2PError) as e:
        E.Simulation(par_list=par, mode=0)
    assert "N_R" in str(e.value)


@pytest.mark.parametrize("exe,mode", [("MoSES_2PDF_landslide", 0), ("MoSES_2PDF_inflow", 1)])
def test_cli_reports_invalid_par_list_values(tmp_path, exe, mode):
    # The executables print the reason and exit 1, before any device work, where the
    # reference would run (or divide by zero) with the values it read.
    import subprocess
    from conftest import ROOT
    path = os.path.join(ROOT, "moses_2pdf_amd", "bin", exe)
    if not os.path.exists(path):
        pytest.skip("executables not built")
    if mode == 0:
        sc = S.Scenario(mode=0, topo=np.zeros((8, 8)), depth=np.ones((8, 8)), cellsize=20.0,
                        total_sim=10.0, each_out=1.0)
    else:
        sc = S.inflow_valley(nx=16, ny=16, nloc=4)
    S.write_case(sc, str(tmp_path))
    par = tmp_path / "par_List"
    lines = par.read_text().split("\n")
    good = list(lines)
    lines[4] = "0.0   # value of N_R"
    par.write_text("\n".join(lines))
    r = subprocess.run([path], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "N_R" in r.stdout
    par.write_text("\n".join(good))
    r = subprocess.run([path, "cfl=fast"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "overlay" in r.stdout
