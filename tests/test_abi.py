"""The C-ABI library loads and exports every entry point include/octofmm.h
declares; argument checks that need no device."""
import ctypes
import os
import re

import pytest

from octotiger_amd import _abi, native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "octofmm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(octofmm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert declared_functions() == sorted(native.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    for fn in declared_functions():
        assert hasattr(lib, fn), fn


def test_abi_version_and_status_strings():
    lib = native.lib()
    assert lib.octofmm_abi_version() == 1
    assert lib.octofmm_status_string(_abi.OK) == b"ok"
    assert lib.octofmm_status_string(_abi.ERR_THETA).startswith(b"invalid theta")


def test_null_arguments_are_rejected_without_a_device():
    lib = native.lib()
    out = ctypes.c_void_p()
    assert lib.octofmm_create(None, ctypes.byref(out)) == _abi.ERR_INVALID_ARGUMENT
    assert lib.octofmm_compute_multipole_interactions(None, None, None, None) == _abi.ERR_INVALID_ARGUMENT
    assert lib.octofmm_compute_monopole_interactions(None, None, 0, None, None) == _abi.ERR_INVALID_ARGUMENT
    assert lib.octofmm_level_plan_execute(None, None) == _abi.ERR_INVALID_ARGUMENT
    assert lib.octofmm_destroy(None) == _abi.OK


def test_struct_layouts_match_header_sizes():
    # 6 pointer-or-int fields, 8-byte aligned
    assert ctypes.sizeof(_abi.Neighbor) == 40
    assert ctypes.sizeof(_abi.Subgrid) == 3 * 8 + 26 * 40 + 8 + 24 + 8
    assert ctypes.sizeof(_abi.PlanStats) == 6 * 8 + 4 * 4
