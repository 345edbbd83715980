"""The C ABI library loads and exports every function include/moses2pdf.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "moses2pdf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(m2p_[a-z_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "m2p_open" in names and "m2p_run" in names and "m2p_step" in names
    assert len(names) >= 16


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "moses_2pdf_amd", "libmoses2pdf.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_all_exports():
    from moses_2pdf_amd import engine
    assert sorted(engine.EXPORTS) == declared()


def test_default_params_are_the_reference_constants():
    from moses_2pdf_amd import default_params
    p = default_params(0)
    assert p.cfl == 0.1 and p.rho_ratio == 1.420 / 2.60 and p.dt_min == 0.001 and p.dt_max == 2.0
    q = default_params(1)
    assert q.dt_min == 0.002
