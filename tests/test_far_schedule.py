"""The far schedule of the direct-staging kernels, checked on the host: each
far pair of the interaction tables is evaluated exactly once, and at
theta = 0.5 the split face images save 12 of the 208 lockstep steps per cell
(tests/cpp/far_schedule_check.cpp compiled with the library's stencil.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "octotiger_amd", "csrc")


def test_far_schedule_covers_every_far_pair_once(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    exe = tmp_path / "far_schedule_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include"),
                    "-I" + CSRC, os.path.join(ROOT, "tests", "cpp", "far_schedule_check.cpp"),
                    os.path.join(CSRC, "stencil.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
    rows = {float(l.split()[1]): l.split() for l in out.stdout.splitlines() if l.startswith("theta")}
    r = rows[0.5]
    assert int(r[3]) == 26 and int(r[5]) == 196 and int(r[7]) == 208 and int(r[11]) == 0
