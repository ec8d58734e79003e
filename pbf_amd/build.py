"""Builds libpbf_amd.so in-tree with hipcc for gfx950 (no GPU needed), and the
C++ program of tests/cpp that drives the classes of include/pbf.hpp over it.

    python -m pbf_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpbf_amd.so")
CPP_DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "sph_cpp.cpp")
CPP_DEMO = os.path.join(ROOT, "tests", "cpp", "sph_cpp")
CLI_SRC = os.path.join(HERE, "cli", "pbf_run.cpp")
CLI = os.path.join(HERE, "pbf_run")
SOURCES = ["pbf_kernels.hip", "pbf_sort.hip", "pbf_selftest.hip", "pbf_batched.hip", "pbf_adjoint.hip",
           "pbf_capi.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Every float expression is rounded as written: no fused multiply-add.
    "-ffp-contract=off",
    # No packed FP32 instructions (v_pk_*_f32) in the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    # Range tests combine booleans with & on purpose (no short-circuit branches).
    "-Wno-bitwise-instead-of-logical",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES if os.path.exists(os.path.join(HERE, "csrc", s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{out}")
    if out.strip() and verbose:
        print(out)
    os.replace(tmp, LIB)
    if os.path.exists(CPP_DEMO_SRC):
        build_cpp_demo()
    build_cli()
    return LIB


def build_cpp_demo() -> str:
    """g++ build of the C++ class demo, linked to the in-tree library by rpath."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", CPP_DEMO + ".tmp",
           CPP_DEMO_SRC, "-L", HERE, "-l:libpbf_amd.so", "-Wl,-rpath,$ORIGIN/../../pbf_amd"]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stdout}")
    os.replace(CPP_DEMO + ".tmp", CPP_DEMO)
    return CPP_DEMO


def build_cli() -> str:
    """g++ build of pbf_run (scene files through include/pbf.hpp), linked to the
    in-tree library by rpath."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           "-o", CLI + ".tmp", CLI_SRC, "-L", HERE, "-l:libpbf_amd.so", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stdout}")
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
