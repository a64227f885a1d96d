"""Build recipe of libhila_wilson.so (gfx950) and its C++ test programs.

    python -m hila_amd.build            # compile everything in-tree
    python -m hila_amd.build --resources  # also print per-kernel VGPR/SGPR/LDS usage

Device code is compiled by hipcc for gfx950 only; packed FP32 instructions are disabled
(-packed-fp32-ops).  The host-only input generator is compiled by g++ with OpenMP.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hila_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "hila_amd", "_build")
LIBDIR = os.path.join(ROOT, "hila_amd", "lib")
LIB = os.path.join(LIBDIR, "libhila_wilson.so")

HIP_SOURCES = ["wl_rb_f16.hip", "wl_rb_f32.hip", "wl_rb_f32r.hip", "wl_rb_f64.hip", "wl_rb_f64r.hip",
               "wl_dslash.hip", "wl_multi.hip",
               "wl_blas.hip", "wl_api.hip", "wl_cg.hip", "wl_cgp.hip", "wl_bcg.hip", "wl_noalloc.hip", "wl_serve.hip", "wl_config.hip", "wl_mre.hip", "wl_force.hip"]
# Per-source device flags.  wl_rb_f16.hip: without the v_dot2c_f32_f16 encoding (dot5-insts,
# which accumulates in place and so needs a zero move to start every FP16 dot-product chain)
# the compiler selects v_dot2_f32_f16, whose accumulator is a separate operand (an inline 0 at
# the start of a chain): 10% fewer VALU instructions in the FP16 hopping kernels.
SOURCE_FLAGS = {"wl_rb_f16.hip": ["-Xclang", "-target-feature", "-Xclang", "-dot5-insts"]}
# C++ programs written against the HILA-style mirror (include/hila_wilson.hpp)
CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", "test_wilson_evenodd.cpp")]
CPP_TEST_BUILD = os.path.join(ROOT, "tests", "cpp", "_build")
HOST_SOURCES = ["wl_gen.cpp"]

ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def hip_flags(resources=False):
    f = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, "-I", CSRC,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
    if resources:
        f += ["-Rpass-analysis=kernel-resource-usage"]
    return f


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = (p.stdout or "") + (p.stderr or "")
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(resources=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([hipcc(), *hip_flags(resources), *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
                     "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        jobs.append(["g++", "-O2", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-I", INCLUDE,
                     "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose or resources:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lgomp", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    os.makedirs(CPP_TEST_BUILD, exist_ok=True)
    for src in CPP_TESTS:
        exe = os.path.join(CPP_TEST_BUILD, os.path.basename(src)[:-4])
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, src, "-o", exe, "-L", LIBDIR,
              "-lhila_wilson", "-Wl,-rpath," + LIBDIR + ",-rpath,$ORIGIN/../../../hila_amd/lib"])
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--resources", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(resources=a.resources, verbose=a.verbose))


if __name__ == "__main__":
    sys.exit(main())
