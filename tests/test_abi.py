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


def test_struct_layouts_match_the_header_compiled_by_gcc(tmp_path):
    """Every ctypes mirror has the size and field offsets gcc gives the C
    structs of include/octofmm.h."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    structs = {"octofmm_neighbor": _abi.Neighbor, "octofmm_subgrid": _abi.Subgrid, "octofmm_config": _abi.Config,
               "octofmm_partial_slot": _abi.PartialSlot, "octofmm_level": _abi.Level,
               "octofmm_plan_stats": _abi.PlanStats, "octofmm_tree_level": _abi.TreeLevel,
               "octofmm_tree": _abi.Tree, "octofmm_tree_stats": _abi.TreeStats}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "octofmm.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
