"""The C ABI library loads and exports every entry point include/bsmps.h declares."""
import ctypes
import os
import re

import pytest

from bosonsupremacy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bsmps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsmps_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_exports_all_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libbsmps.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_error_paths_without_device():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libbsmps.so not built")
    lib = _lib.load()
    assert lib.bsmps_version() >= 1
    assert lib.bsmps_set_tuning(None, 4, 64) == _lib.BSMPS_ERR_INVALID
    assert b"set_tuning" in lib.bsmps_last_error()
    assert lib.bsmps_synchronize(None) == _lib.BSMPS_ERR_INVALID
    assert lib.bsmps_create(0, None, None) == _lib.BSMPS_ERR_INVALID
    assert lib.bsmps_set_split(None, 12, 1) == _lib.BSMPS_ERR_INVALID
    assert lib.bsmps_set_split_budget(None, 1 << 30) == _lib.BSMPS_ERR_INVALID
    assert lib.bsmps_set_contract_form(None, 0, 1) == _lib.BSMPS_ERR_INVALID
    with pytest.raises(_lib.BsmpsError):
        _lib.check(lib.bsmps_synchronize(None))


def test_column_mapped_contractions_refuse_bad_arguments_without_device():
    """The _cols contractions refuse a call without a context before any device work and name
    themselves in the error (their other refusals need a context: tests/test_gpu_sampling.py)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libbsmps.so not built")
    lib = _lib.load()
    p = ctypes.c_void_p(8)          # never dereferenced: every call below is refused first
    for name, ops in (("bsmps_site_contract_cols", (p, p, p, p)), ("bsmps_site_contract_c64_cols", (p, p, p, p)),
                      ("bsmps_site_contract_bf16x6_cols", (p, p, p)),
                      ("bsmps_site_contract_fp16x3_cols", (p, p, p, p, p))):
        fn = getattr(lib, name)
        assert fn(None, 10, 4, 3, 2, *ops, 6, p) == _lib.BSMPS_ERR_INVALID
        assert name.encode() in lib.bsmps_last_error()
