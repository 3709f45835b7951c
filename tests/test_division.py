"""The tracer's quotient (p - o) / h without the division sequence, its
remainder from the cell subscript and the reference weight fmod(d, h) / h,
against IEEE division and fmodf on the host.

The device holds 1/h as y + yl (y = RN(1/h)), takes q0 = RN(a y + RN(a yl)),
within one ulp of a/h, and one FMA correction, RN(a/h) by Markstein's
theorem for a in {0} or [2^-40, 2^40] and h in [2^-20, 2^20] (outside, the
device divides).  The remainder of a located point is d - m h with m =
floor(q), plus h when q rounded up to m.  These checks cover whole binades
of a for the spacings the workloads use and strided sweeps of the full range
for awkward spacings (all-ones significands, powers of two)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def div_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("div") / "libdiv_check.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "cpp", "div_check.c"), "-lm"], check=True)
    L = C.CDLL(str(out))
    for f in (L.quotient_mismatches, L.fmod_mismatches, L.weight_mismatches):
        f.restype = C.c_uint64
        f.argtypes = [C.c_float, C.c_uint32, C.c_uint32, C.c_uint32]
    L.edge_mismatches.restype = C.c_uint64
    L.edge_mismatches.argtypes = [C.c_float]
    return L


def bits(x):
    return int(np.float32(x).view(np.uint32))


WORKLOAD_SPACINGS = [np.float32(2 * math.pi / 511), np.float32(2 * math.pi / 63), np.float32(2 * math.pi / 255),
                     np.float32(2 * math.pi / 1023), np.float32(1 / 47)]
AWKWARD_SPACINGS = [np.float32(1.0), np.float32(0.5), np.float32(3.0), np.float32(1 / 3),
                    np.float32(0.1), np.uint32(0x3F7FFFFF).view(np.float32), np.uint32(0x3FFFFFFF).view(np.float32),
                    np.float32(2.0 ** -20), np.float32(2.0 ** 20), np.uint32(0x49FFFFFF).view(np.float32)]


@pytest.mark.parametrize("h", WORKLOAD_SPACINGS, ids=lambda h: "%.6g" % h)
def test_quotient_whole_binades(div_lib, h):
    # every float a in [2^-3, 2^3): the block-local coordinates of a 128-cell block
    assert div_lib.quotient_mismatches(float(h), bits(0.125), bits(8.0), 1) == 0


@pytest.mark.parametrize("h", WORKLOAD_SPACINGS + AWKWARD_SPACINGS, ids=lambda h: "%.6g" % h)
def test_quotient_strided_full_range(div_lib, h):
    # a over [2^-40, 2^40] on a stride that visits every significand pattern class
    assert div_lib.quotient_mismatches(float(h), bits(2.0 ** -40), bits(2.0 ** 40), 977) == 0
    assert div_lib.quotient_mismatches(float(h), 0, 1, 1) == 0  # +-0 keeps its sign


@pytest.mark.parametrize("h", WORKLOAD_SPACINGS + AWKWARD_SPACINGS, ids=lambda h: "%.6g" % h)
def test_weight_from_subscript(div_lib, h):
    # located points d >= 0 with d / h < 2^22: whole binade [2^-3, 2^3) and a strided sweep down to 2^-40
    hi = min(bits(float(h) * 4194300.0), bits(2.0 ** 40))
    assert div_lib.weight_mismatches(float(h), bits(0.125), bits(8.0), 1) == 0
    assert div_lib.weight_mismatches(float(h), bits(2.0 ** -40), hi, 331) == 0
    assert div_lib.edge_mismatches(float(h)) == 0


@pytest.mark.parametrize("h", WORKLOAD_SPACINGS + AWKWARD_SPACINGS[:6], ids=lambda h: "%.6g" % h)
def test_fmod_branch_free(div_lib, h):
    hi = min(bits(float(h) * 4194300.0), bits(2.0 ** 40))
    assert div_lib.fmod_mismatches(float(h), 1, hi, 331) == 0
    assert div_lib.fmod_mismatches(float(h), bits(0.125), bits(8.0), 7) == 0
