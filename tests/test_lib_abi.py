"""The C-ABI library loads and exports every entry point the public header
declares (no GPU needed: no compute call is made)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qbx_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*\s*|int\s+|long\s+)(qbx_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 35
    for s in ["qbx_tree_build", "qbx_traverse", "qbx_target_assoc", "qbx_geo_build",
              "qbx_form_global_qbx_locals", "qbx_translate_box_multipoles_to_qbx_local",
              "qbx_translate_box_local_to_qbx_local", "qbx_eval_qbx_expansions",
              "qbx_drive_fmm"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    from pytential_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libqbx_mi355x.so not built (run __graft_entry__.build())")
    lib = _lib.lib()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    # the Python binding covers exactly the declared surface
    assert set(declared_symbols()) == set(_lib.exported_symbols())


def test_library_reports_errors_without_gpu_work():
    import ctypes as ct

    from pytential_amd import _lib
    lib = _lib.lib()
    kd = _lib.KernelDesc(7, 0.0, 0, 4, 1, None)
    assert lib.qbx_ncoeff_qbx(ct.byref(kd)) == -1
    assert b"unsupported kernel" in lib.qbx_last_error()
    lo = (ct.c_int * 2)(10, 12)
    kd = _lib.KernelDesc(_lib.HELMHOLTZ2D, 20.0, _lib.SOURCE_D, 5, 2, lo)
    assert lib.qbx_ncoeff_qbx(ct.byref(kd)) == 11
    kd = _lib.KernelDesc(_lib.LAPLACE2D, 0.0, _lib.SOURCE_S, 4, 2, lo)
    assert lib.qbx_ncoeff_qbx(ct.byref(kd)) == 5
    # a null tree is rejected without touching the device
    assert lib.qbx_expansion_size(ct.byref(kd), None) == -1


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    from pytential_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libqbx_mi355x.so")
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()
