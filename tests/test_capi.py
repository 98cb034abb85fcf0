"""The C-ABI library loads and exports every entry point include/atm_mi355x.h declares;
without a GPU every entry point that needs one fails with a status, not a crash."""
import ctypes
import os
import re

import pytest

from atmmetaforce import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "atm_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(atm_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_and_abi():
    lib = _capi.load()
    assert lib.atm_abi_version() == 6
    assert b"atm-mi355x" in lib.atm_version()


def test_errors_are_statuses():
    lib = _capi.load()
    # null context: every context entry point reports an invalid argument
    assert lib.atm_eval(None, 0) == _capi.ATM_ERR_INVALID_ARGUMENT
    assert b"null context" in lib.atm_last_error()
    r = _capi.Result()
    assert lib.atm_get_results(None, 0, ctypes.byref(r)) == _capi.ATM_ERR_INVALID_ARGUMENT
    assert lib.atm_destroy(None) == _capi.ATM_OK
    # invalid system descriptions are rejected before any device call
    d = (_capi.SystemDesc * 1)()
    d[0].num_atoms = 10
    d[0].box[0] = d[0].box[1] = d[0].box[2] = 1.0
    d[0].cutoff = 0.6  # more than half the box
    d[0].rf_dielectric = 78.5
    h = ctypes.c_void_p()
    rc = lib.atm_create(d, 1, 0, ctypes.byref(h))
    assert rc == _capi.ATM_ERR_INVALID_ARGUMENT
    assert b"half the periodic box" in lib.atm_last_error()
    d[0].num_atoms = 0
    assert lib.atm_create(d, 1, 0, ctypes.byref(h)) == _capi.ATM_ERR_INVALID_ARGUMENT


def test_no_gpu_means_clean_failure():
    n = _capi.device_count()
    if n > 0:
        pytest.skip("a GPU is visible")
    lib = _capi.load()
    d = (_capi.SystemDesc * 1)()
    d[0].num_atoms = 10
    d[0].box[0] = d[0].box[1] = d[0].box[2] = 3.0
    d[0].cutoff = 1.0
    d[0].rf_dielectric = 78.5
    h = ctypes.c_void_p()
    rc = lib.atm_create(d, 1, 0, ctypes.byref(h))
    assert rc in (_capi.ATM_ERR_INVALID_ARGUMENT, _capi.ATM_ERR_DEVICE)
    assert not h.value
