"""The C ABI library loads and exports every function include/chemsim.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import re

import pytest

from chemsim_amd import capi


def header_functions():
    text = open(capi.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cs_\w+)\s*\(", text, flags=re.M)))


def test_header_parsed():
    names = header_functions()
    assert "cs_lattice_step" in names and "cs_last_error" in names
    assert len(names) >= 40


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    declared = set(header_functions()) - {"cs_last_error", "cs_abi_version"}
    assert declared <= set(capi.SIGNATURES), sorted(declared - set(capi.SIGNATURES))


def test_abi_version_and_error_path():
    assert capi.lib.cs_abi_version() == 1
    h = C.c_void_p()
    rc = capi.lib.cs_ctx_create(0, C.byref(h))
    if rc == capi.CS_OK:          # a GPU is present
        assert capi.lib.cs_ctx_destroy(h) == capi.CS_OK
    else:
        assert rc in (capi.CS_ERR_HIP, capi.CS_ERR_ARG)
        assert capi.last_error()
    # argument errors are reported, not crashed on
    assert capi.lib.cs_lattice_step(None, 1) == capi.CS_ERR_ARG
    assert "null" in capi.last_error()
    assert capi.lib.cs_set_cells_per_thread(3) == capi.CS_ERR_ARG
    assert capi.lib.cs_set_cells_per_thread(-3) == capi.CS_ERR_ARG
    assert capi.lib.cs_set_cells_per_thread(capi.CS_STEP_VECTOR_WIDE) == capi.CS_OK
    assert capi.lib.cs_set_cells_per_thread(capi.CS_STEP_VECTOR) == capi.CS_OK
    assert capi.lib.cs_set_cells_per_thread(0) == capi.CS_OK
    for k in (2, 3, 6, -2):
        assert capi.lib.cs_set_temporal_steps(k) == capi.CS_ERR_ARG
    for k in (0, 4, 8, capi.CS_TEMPORAL_AUTO):
        assert capi.lib.cs_set_temporal_steps(k) == capi.CS_OK
    k = C.c_int()
    assert capi.lib.cs_lattice_temporal_steps(None, C.byref(k)) == capi.CS_ERR_ARG
    n = C.c_uint64()
    assert capi.lib.cs_check_reciprocal(None, C.byref(n)) == capi.CS_ERR_ARG
