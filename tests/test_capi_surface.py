"""The C-ABI library loads and exports every function include/dense_capi.h
declares, and the Fortran binding names the same symbols (no compute calls:
these run without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT, dense

HEADER = os.path.join(ROOT, "include", "dense_capi.h")
FMOD = os.path.join(ROOT, "2dense_amd", "fortran", "dense_capi_mod.f90")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dense_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dense_create", "dense_destroy", "dense_set_state", "dense_get_state", "dense_step",
                 "dense_run", "dense_update_params", "dense_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dense().LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_fortran_binding_names_exist_in_header():
    f = open(FMOD).read()
    bound = set(re.findall(r'bind\(C,\s*name="(dense_[a-z_0-9]+)"\)', f))
    assert bound and bound <= set(declared())


def test_cfg_struct_layout_matches_header():
    """dense_cfg offsets as seen by ctypes (the Python binding) equal a C
    compiler's; the Fortran bind(C) type follows the same C rules."""
    D = dense()
    assert ctypes.sizeof(D.Cfg) == 240
    assert D.Cfg.bc_state.offset == 104 and D.Cfg.tile.offset == 232


def test_ic_struct_layout_matches_header():
    import importlib
    B = importlib.import_module("2dense_amd.batch")
    assert ctypes.sizeof(B.IcDesc) == 88
    assert B.IcDesc.eps.offset == 8 and B.IcDesc.alpha.offset == 80
