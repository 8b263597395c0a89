"""The sparse plan builder (spicier_amd/csrc/sparse_plan.hip) on the CPU:
built with g++ into a checker (tests/native/plan_check.cpp) that verifies,
on MNA grids from 8x8 to 40x40, that every solve descriptor the DC kernels
read equals the row tables it summarises, that the LDS-staged prefix of the
plan ends where the descriptors begin, and that the dense root block is one
column per level and closed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "spicier_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_plan_descriptors_and_root_block(tmp_path):
    exe = str(tmp_path / "plan_check")
    cmd = ["g++", "-O1", "-std=c++17", "-D__host__=", "-D__device__=", f"-I{CSRC}",
           os.path.join(HERE, "native", "plan_check.cpp"), "-x", "c++",
           os.path.join(CSRC, "sparse_plan.hip"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "8", "16", "24", "32", "40"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 5 and all(l.endswith("ok") for l in lines), r.stdout
    assert "grid 32: n 1028 nnz_lu 23668 levels 118 tail 48" in r.stdout
