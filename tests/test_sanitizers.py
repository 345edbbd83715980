"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only; GPU code
is not built with sanitizers): the CPU oracle driven through every entry point on a small
landslide and a small inflow case (tests/oracle_sanitize.cpp), and the product's file
readers and writers (m2p_io.cpp) on the reference's own datasets and on malformed input
(tests/io_check.cpp).  Any out-of-bounds access, use after free, leak or undefined
operation (signed overflow, invalid shift, misaligned access, out-of-range conversion)
aborts the program, and the test fails."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from util import REFDATA

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", OMP_NUM_THREADS="1",
           UBSAN_OPTIONS="print_stacktrace=1")


def _build(tmp, name, srcs, extra=()):
    exe = str(tmp / name)
    subprocess.run(["g++", "-O1", "-std=c++17", *SAN, *extra, "-o", exe, *srcs], check=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_oracle_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_sanitize",
                 [os.path.join(ROOT, "tests", "oracle_sanitize.cpp"),
                  os.path.join(ROOT, "oracle", "m2p_oracle.cpp")],
                 ["-ffp-contract=off", "-fopenmp"])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mode 0:" in r.stdout and "mode 1:" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_file_io_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "io_check", [os.path.join(ROOT, "tests", "io_check.cpp"),
                                        os.path.join(ROOT, "moses_2pdf_amd", "csrc", "m2p_io.cpp")])

    def run(*args, cwd=None):
        r = subprocess.run([exe, *args], capture_output=True, text=True, env=ENV, cwd=cwd, timeout=120)
        assert "Sanitizer" not in r.stderr, r.stderr
        return r.returncode

    assert run("par", "0", os.path.join(REFDATA, "par_List.landslide")) == 0
    assert run("par", "1", os.path.join(REFDATA, "par_List.inflow")) == 0
    for name, as_float in (("shiaolin_topo_20m.asc", "1"), ("shiaolin_init20m.asc", "0"),
                           ("00_DebrisFlow_dem_20m.asc", "1")):
        assert run("asc", os.path.join(REFDATA, name), as_float) == 0
    # malformed inputs: errors, not crashes
    bad = tmp_path / "bad.asc"
    bad.write_text("ncols 4\nnrows 3\nxllcorner 0\nyllcorner 0\ncellsize 20\n1 2 3\n")
    assert run("asc", str(bad), "1") != 0
    (tmp_path / "short_par").write_text("10\n1\n")
    assert run("par", "0", str(tmp_path / "short_par")) != 0
    assert run("par", "0", str(tmp_path / "missing")) != 0
    out = tmp_path / "out"
    out.mkdir()
    assert run("write", str(out)) == 0
