"""Build libpa_gpu.so (HIP for gfx950) in-tree.

Every translation unit is compiled with contraction off (each float
multiply and add rounded separately, as the reference's SSE arithmetic
is), IEEE division and square root, and packed-FP32 instruction selection
disabled.  Usage: ``python -m pars_amd.build`` or ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libpa_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PA_GPU_ARCH", "gfx950")

SOURCES = ["ctx.hip", "field.hip", "trace.hip", "trace_coop.hip", "trace_quad.hip", "compact.hip", "color.hip",
           "round.hip", "batch.hip", "ftle.hip", "noalloc.hip", "handover.hip", "io.hip", "graph.hip", "serve.hip"]

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cpp_demo(verbose: bool = False) -> str:
    """Compile tests/cpp/stages_demo against the C++ stage mirror and libpa_gpu.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "stages_demo.cpp")
    exe = os.path.join(root, "tests", "cpp", "stages_demo")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include"), "-I",
           os.path.join(HERE, "include"), src, "-o", exe, "-L", HERE, "-lpa_gpu", "-Wl,-rpath,$ORIGIN/../../pars_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("stages_demo build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_demo(verbose=True)
    sys.exit(0)
