"""CPU test of the site-tiled device layout (tile_elem, hila_amd/csrc/wl_store.h): the host
program tests/cpp/test_tile_layout.hip is compiled host-only with hipcc and run; it checks
that every block shape the kernels use maps (entry, site) one-to-one onto the block, that a
16-byte vector is P consecutive, P-aligned elements, and that a tile's 64 sites read one
contiguous run per vector."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.mark.skipif(HIPCC is None, reason="needs hipcc")
def test_tile_layout_is_a_bijection_with_contiguous_vectors(tmp_path):
    exe = tmp_path / "test_tile_layout"
    subprocess.run([HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "hila_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_tile_layout.hip"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "layout ok" in p.stdout
