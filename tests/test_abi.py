"""The C ABI library: builds, loads, exports every entry point that
include/cap_amd.h declares, and refuses cleanly without a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cap_amd.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(cap_\w+)\(", txt, re.M)))


def test_header_declares_the_binding_list(cap):
    assert declared() == sorted(cap.EXPORTS)


def test_library_exports_every_declared_symbol(cap):
    lib = C.CDLL(cap.host.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_every_entry_point_has_a_prototype(cap):
    """ctypes passes a Python int without a prototype as a 32-bit C int, which
    would cut a 64-bit device or host address: every entry point the binding
    calls must have its argument types declared"""
    lib = cap.load()
    missing = [s for s in cap.EXPORTS if getattr(lib, s).argtypes is None]
    assert not missing, missing


def test_fortran_binding_names_match(cap):
    fpath = os.path.join(ROOT, "cappuccino-ghostcells-mpi_amd", "fortran", "cap_hip.f90")
    if not os.path.exists(fpath):
        pytest.skip("no Fortran binding module")
    txt = open(fpath).read().lower()
    for s in declared():
        if s in ("cap_last_error", "cap_time_kernels", "cap_build_lognormal"):
            continue
        assert f"bind(c, name='{s}')" in txt, s


def test_no_device_is_an_error_not_a_crash(cap):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(cap.CapError):
        cap.Context(12, 10, 9)
