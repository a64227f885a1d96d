"""The C++ mirror of HILA's call surface (include/hila_wilson.hpp), exercised by a C++ program
that restates the Wilson even-odd block of HILA's applications/test_cases/test_CG.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_wilson_evenodd")


def test_cpp_program_is_built():
    assert os.path.exists(EXE), "build with python -m hila_amd.build"


@pytest.mark.gpu
def test_cpp_test_cg_wilson_evenodd_block():
    print("running", EXE, flush=True)
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASS" in p.stdout
