"""The C-ABI library loads and exports every symbol the header declares (no
compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tbplas_amd.h")
LIB = os.path.join(ROOT, "tbplas_amd", "_build", "libtbplas_amd.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tbg_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from tbplas_amd import build
        build.build()
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    names = declared()
    for must in ("tbg_corr_dos", "tbg_corr_ac", "tbg_corr_dc", "tbg_ham_dxy_count",
                 "tbg_ham_dxy_fill", "tbg_ham_csr_count", "tbg_ham_csr_fill",
                 "tbg_ham_create", "tbg_corr_dos_batch", "tbg_dos_begin"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from tbplas_amd import _lib
    assert set(declared()) <= set(_lib.exported_symbols())


def test_version_and_error_string_without_gpu(lib):
    lib.tbg_version.restype = ctypes.c_int
    assert lib.tbg_version() >= 10000
    lib.tbg_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.tbg_last_error(), bytes)


def test_null_handle_is_an_error_not_a_crash(lib):
    # argument checks run before any device call
    lib.tbg_ham_info.restype = ctypes.c_int
    assert lib.tbg_ham_info(None, None, None, None, None) == 2
    lib.tbg_last_error.restype = ctypes.c_char_p
    assert b"null" in lib.tbg_last_error()
