"""The C-ABI library loads and exports every entry point include/ipm_hip.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ipm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ipm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for name in ("ipm_create", "ipm_move_es", "ipm_absorb", "ipm_p2c", "ipm_field_sum", "ipm_mcc", "ipm_step",
                 "ipm_step_graph", "ipm_sort", "ipm_set_fields", "ipm_upload_species", "ipm_download_species"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from ipm2d import _native
    lib = _native.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from ipm2d import _native
    declared = set(declared_symbols()) - {"ipm_last_error"}
    assert declared <= set(_native.SIGNATURES), sorted(declared - set(_native.SIGNATURES))


def test_fortran_module_binds_header():
    """The iso_c_binding module the Fortran host uses names every entry
    point of the header (the reference's side of the boundary is Fortran)."""
    mod = open(os.path.join(ROOT, "2dmatching-cf4_amd", "fortran", "ipm_hip_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(ipm_[a-z_0-9]+)"\)', mod))
    missing = sorted(set(declared_symbols()) - bound)
    assert not missing, f"not bound in ipm_hip_mod.f90: {missing}"


def test_create_without_gpu_fails_cleanly():
    """Without a device the library reports IPM_ENODEV instead of crashing."""
    from conftest import gpu_available
    if gpu_available():
        return
    from ipm2d import _native
    lib = _native.load()
    h = ctypes.c_void_p()
    rc = lib.ipm_create(ctypes.byref(h), 17, 9, 1e-3, 5e-11, 3, 1)
    assert rc == _native.IPM_ENODEV
    assert b"device" in lib.ipm_last_error()
