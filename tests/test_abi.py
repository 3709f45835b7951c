"""C ABI surface checks that need no GPU: the library loads, exports every
entry point include/pa_gpu.h declares, and the shared structs have the
reference's layout."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pa_gpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pa_gpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import pars_amd

    lib = pars_amd.lib()
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", pars_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pa_gpu_[a-z0-9_]+)", out))
    assert exported == set(names), (exported ^ set(names))


def test_struct_layouts():
    import pars_amd

    assert pars_amd.PARTICLE_DTYPE.itemsize == 48  # pa::particle with Vector4f alignment
    assert C.sizeof(pars_amd.RoundInfo) == 40
    assert C.sizeof(pars_amd._Handoff) == 8 + 8 + 8 + 12 * 8
    assert C.sizeof(pars_amd._PoolC) == 8 * 8 + 8  # eight pointers and n_requests (padded)
    assert pars_amd._PoolC.n_requests.offset == 64


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include <stddef.h>\n#include "pa_gpu.h"\n_Static_assert(sizeof(pa_particle) == 48, "particle");\n'
                   "_Static_assert(sizeof(pa_round_info) == 40, \"round\");\n"
                   "_Static_assert(sizeof(pa_pool) == 72 && offsetof(pa_pool, n_requests) == 64, \"pool\");\n"
                   "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                    "-o", str(tmp_path / "t.o")], check=True)
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "pa_amd/stages.hpp"\nint main(){return 0;}\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "pars_amd", "include"), "-fsyntax-only", str(cpp)], check=True)


def test_no_gpu_calls_fail_cleanly():
    """Without a GPU, context creation reports an error instead of crashing."""
    import torch

    import pars_amd

    if torch.cuda.is_available():
        pytest.skip("GPU present; covered by the gpu tests")
    with pytest.raises(pars_amd.PaError):
        pars_amd.Context(0)


def test_seed_count_host_only():
    import pars_amd

    h = np.float32(2 * np.pi / 63)
    n, dims = pars_amd.seed_generator.count([np.float32(64) * h] * 3, [np.float32(4) * h] * 3)
    assert n == 4096 and dims == (16, 16, 16)
    h = np.float32(2 * np.pi / 511)
    n, dims = pars_amd.seed_generator.count([np.float32(128) * h] * 3, [np.float32(2) * h] * 3)
    assert n == 262144 and dims == (64, 64, 64)
