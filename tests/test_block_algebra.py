"""CPU test of the block-CG small-matrix algebra (hila_amd/csrc/wl_small.h): the C++ program
tests/cpp/test_small_algebra.cpp is compiled with g++ and run; it checks the Jacobi
eigendecomposition, SVQB and the pseudo-inverse against their defining identities, with a
repeated column and a zero matrix among the cases."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_small_matrix_algebra(tmp_path):
    exe = tmp_path / "test_small_algebra"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "hila_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_small_algebra.cpp"), "-o", str(exe)], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all ok" in p.stdout
