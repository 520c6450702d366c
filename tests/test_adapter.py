"""The header-only C++ adapter (include/octofmm_interface.hpp) keeps the
reference's interface signatures: it builds against Octo-Tiger-shaped types
and, on a GPU, gives the oracle's results."""
import os
import subprocess

import numpy as np
import pytest

import fmm_cases
import oracle
from metrics import normwise
from octotiger_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp")
LIBDIR = os.path.join(ROOT, "octotiger_amd")


def build(tmp_path):
    exe = os.path.join(str(tmp_path), "adapter_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC,
           "-L" + LIBDIR, "-loctofmm", "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_adapter_compiles_against_octotiger_shaped_types(tmp_path):
    assert os.path.exists(build(tmp_path))


def write_case(path, name, t, spec=None):
    """The case's 27 slots; the neighbours in spec (raw direction -> cell
    list) hand over only those cells, listed by the stand-in grid's
    get_ilist_n_bnd."""
    lv = fmm_cases.build_level(name)
    kind = dict(fmm_cases.CASES)[name]
    cmp_ = any(lv.neighbors[13, r] >= 0 and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_NONLEAF
               for r in range(27) if r != 13)
    spec = spec or {}
    with open(path, "wb") as f:
        np.array([1 if kind == "monopole" else 0, t, 1 if kind == "root" else 0, 1 if cmp_ else 0],
                 np.int32).tofile(f)
        np.array([lv.dx, *lv.xbase[13]], np.float64).tofile(f)
        for raw in range(27):
            n = int(lv.neighbors[13, raw])
            k = -1 if n < 0 else int(lv.kind[n])
            cells = np.asarray(spec.get(raw, np.arange(512)), np.int32)
            np.array([k, cells.size], np.int32).tofile(f)
            cells.tofile(f)
            s = max(n, 0)
            for a in (lv.M[s], lv.com[s], lv.mon[s]):
                np.ascontiguousarray(a[cells], np.float64).tofile(f)
    return lv, kind, cmp_


@pytest.mark.gpu
@pytest.mark.parametrize("name,pattern", [("mixed_halo", None), ("c1_root", None), ("leaf_p2m", None),
                                          ("mixed_halo", "subset"), ("c1_nonroot", "repeats"),
                                          ("leaf_p2m", "subset")])
def test_adapter_on_gpu_matches_oracle(tmp_path, name, pattern):
    """Octo-Tiger-shaped calls through the adapter, full halos and partial
    halos (the adapter reads the cells of a partial halo from the grid's
    get_ilist_n_bnd list, as update_input does), against the oracle."""
    exe = build(tmp_path)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    spec = fmm_cases.partial_specs(name)[pattern] if pattern else None
    lv, kind, cmp_ = write_case(inp, name, _abi.RHO, spec)
    print(f"adapter case {name} partial {pattern}", flush=True)
    subprocess.run([exe, inp, out], check=True, timeout=120)
    res = np.fromfile(out, np.float64)
    L, Lc = res[: 512 * 20].reshape(512, 20), res[512 * 20:].reshape(512, 3)
    if spec:
        sg, keep = fmm_cases.partial_subgrid(lv, spec)
    else:
        sg, keep, _, _ = fmm_cases.subgrid(name, _abi.RHO)
    if kind == "monopole":
        Lr, Lcr = oracle.monopole_interactions(sg, cmp_)
    else:
        Lr, Lcr = oracle.multipole_interactions(sg)
    assert normwise(L, Lr) <= 1e-12 and normwise(Lc, Lcr) <= 1e-12
