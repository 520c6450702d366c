"""Checks on the gfx950 code the compiler makes from the kernels (CPU only:
hipcc cross-compiles without a GPU).

The staged level kernels copy partner records from global memory straight
into LDS (global_load_lds_dwordx4); the LDS destination of such a load is M0,
one value for the whole wave.  If the compiler merges two of these loads with
different destinations from the two sides of a branch, the destination becomes
a per-lane value and is then read from lane 0 (v_readfirstlane_b32 into M0):
every other lane's 16 bytes land in the wrong place and the results are
garbage.  That happened once; this test keeps it from coming back unseen.
"""
import os
import re
import shutil
import subprocess

import pytest

from octotiger_amd import build as B


def _device_asm(tmp_path):
    hipcc = B.hipcc()
    if not os.path.exists(hipcc) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    out = tmp_path / "octofmm.s"
    inc = ["-I" + os.path.join(B.ROOT, "include"), "-I" + B.CSRC]
    cmd = [hipcc] + B.HIP_FLAGS + inc + ["-x", "hip", "--cuda-device-only", "-S",
                                        os.path.join(B.CSRC, "octofmm.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text().splitlines()


def test_lds_load_destinations_are_wave_uniform(tmp_path):
    lines = _device_asm(tmp_path)
    func, from_lane = None, {}
    bad, nloads = [], 0
    for ln in lines:
        m = re.match(r"^(\w+):", ln)
        if m and not ln.startswith("."):
            func, from_lane = m.group(1), {}
            continue
        t = ln.strip()
        if "global_load_lds" in t:
            nloads += 1
        m = re.match(r"v_readfirstlane_b32 (s\d+),", t)
        if m:
            from_lane[m.group(1)] = True
            continue
        m = re.match(r"s_(?:mov_b32|add_i32|add_u32) m0, (s\d+)", t)
        if m and from_lane.get(m.group(1)):
            bad.append(f"{func}: {t}")
        # any other write of an SGPR ends its lane-0 provenance
        m = re.match(r"s_\w+ (s\d+),", t)
        if m:
            from_lane.pop(m.group(1), None)
    assert nloads > 0, "no direct-to-LDS loads found: the check is not looking at the right code"
    assert not bad, "M0 taken from a per-lane value:\n" + "\n".join(bad[:10])
