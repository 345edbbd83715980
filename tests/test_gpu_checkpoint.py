"""Checkpoint / resume: a context restored from a checkpoint continues bit for bit, in
both precisions and both modes, and a run split into segments with a checkpoint and a
fresh process-level context in between writes the same result2Pb/ files as one run."""
import os

import numpy as np
import pytest

from moses_2pdf_amd import M2PError, Simulation
from moses_2pdf_amd import scenarios as S
from oracle.oracle import f32
from util import df004

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,precision", [("fractal", "fp64"), ("fractal", "fp32"),
                                            ("df004", "fp64"), ("plane", "fp32")])
def test_restored_context_continues_bit_for_bit(tmp_path, case, precision):
    sc = {"fractal": lambda: S.landslide_fractal(n=64, seed=7), "df004": df004,
          "plane": lambda: S.landslide_plane(n=48)}[case]()
    ck = str(tmp_path / "state.m2p")
    with Simulation(sc, precision=precision) as a:
        a.set_stop(vstop=0.01)
        a.step(25)
        a.checkpoint(ck)
        a.step(15)
        ua, (ta, na), sa = a.state(), a.time_steps(), a.status()
    with Simulation(sc, precision=precision) as b:
        b.restore(ck)
        assert b.time_steps()[1] == 25
        b.step(15)
        assert np.array_equal(b.state(), ua)
        assert b.time_steps() == (ta, na) and b.status() == sa


def test_segmented_run_writes_the_same_files(tmp_path):
    sc = S.landslide_plane(n=40)
    sc.topo = f32(sc.topo)
    sc.total_sim, sc.each_out = 6.0, 2.0
    root = str(tmp_path / "case")
    par = S.write_case(sc, root)
    one, seg = os.path.join(root, "one"), os.path.join(root, "seg")
    os.makedirs(one)
    os.makedirs(seg)
    ck = os.path.join(root, "run.m2p")
    cwd = os.getcwd()
    os.chdir(root)   # the par_List names its files relative to the working directory
    try:
        with Simulation(par_list=par, mode=0) as s:
            s.run(one + "/")
        finished, segments = False, 0
        while not finished:
            with Simulation(par_list=par, mode=0) as s:
                if segments:
                    s.restore(ck)
                finished = s.run_steps(seg + "/", 37)
                s.checkpoint(ck)
            segments += 1
    finally:
        os.chdir(cwd)
    assert segments >= 3
    names = sorted(os.listdir(one))
    assert names == sorted(os.listdir(seg)) and len(names) >= 6
    for n in names:
        assert open(os.path.join(one, n)).read() == open(os.path.join(seg, n)).read(), n


def test_restore_rejects_mismatch_and_corruption(tmp_path):
    ck = str(tmp_path / "a.m2p")
    with Simulation(S.landslide_plane(n=40)) as a:
        a.step(3)
        a.checkpoint(ck)
    with Simulation(S.landslide_plane(n=48)) as b:
        with pytest.raises(M2PError, match="does not match"):
            b.restore(ck)
    with Simulation(S.landslide_plane(n=40), precision="fp32") as b:
        with pytest.raises(M2PError, match="does not match"):
            b.restore(ck)
    # same grid size, but another DEM, another friction angle, another CFL number
    other_dem = S.landslide_plane(n=40)
    other_dem.topo = other_dem.topo + 1.0
    other_par = S.landslide_plane(n=40)
    other_par.delta0 += 1.0
    other_cfl = S.landslide_plane(n=40)
    other_cfl.cfl = 0.2
    for sc in (other_dem, other_par, other_cfl):
        with Simulation(sc) as b:
            with pytest.raises(M2PError, match="topography, parameters or inflow"):
                b.restore(ck)
    with Simulation(S.landslide_plane(n=40)) as b:   # the same scenario restores
        b.restore(ck)
        assert b.time_steps()[1] == 3
    data = bytearray(open(ck, "rb").read())
    data[len(data) // 2] ^= 0x40
    bad = str(tmp_path / "bad.m2p")
    open(bad, "wb").write(bytes(data))
    with Simulation(S.landslide_plane(n=40)) as b:
        with pytest.raises(M2PError, match="checksum"):
            b.restore(bad)
        with pytest.raises(M2PError, match="truncated"):
            open(bad, "wb").write(bytes(data[:100]))
            b.restore(bad)
