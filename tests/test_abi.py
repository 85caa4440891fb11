"""The C ABI without a GPU: the library loads, exports every entry point
declared in include/vtkgpu.h, and reports errors instead of computing when
it has not been initialised."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vtkgpu.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vk_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_python_export_list():
    from vtkmakecompare_amd._lib import EXPORTS
    assert sorted(EXPORTS) == declared_symbols()


def test_library_exports_every_symbol():
    from vtkmakecompare_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_calls_before_init_fail_loudly():
    from vtkmakecompare_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    n = ctypes.c_int(8)
    rc = L.vk_set_bc(None, None, None, None, ctypes.byref(n), ctypes.byref(n), ctypes.byref(n),
                     None)
    assert rc == 1  # VK_ERR_NOT_INITIALISED
    assert b"vk_init" in L.vk_last_error()


def test_fortran_host_module_builds():
    host = os.path.join(ROOT, "vtkmakecompare_amd", "host")
    if not os.path.exists(os.path.join(host, "Makefile")):
        pytest.skip("no Fortran host")
    exe = os.path.join(host, "vk_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built")
    assert os.access(exe, os.X_OK)
