"""The w component of the frozen-derivative steps: the device adds one
precomputed signed zero (or NaN) instead of the four products a_i * 0 that
the reference's vector4 state sums (particle_tracer.cpp:315-324, odeint's
final stage); checked bit for bit on the host over signed zeros, subnormals,
infinities and NaN of w and of the step size."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def wlib(tmp_path_factory):
    out = tmp_path_factory.mktemp("wstep") / "libwstep_check.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "cpp", "wstep_check.c"), "-lm"], check=True)
    L = C.CDLL(str(out))
    L.wstep_mismatches.restype = C.c_uint64
    L.wstep_mismatches.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    return L


def test_w_step_shortcut(wlib):
    sub = np.uint32(1).view(np.float32)
    ws = np.array([0.0, -0.0, 1.0, -1.0, sub, -sub, 1e-30, -3.5, np.inf, -np.inf, np.nan, 1e38], np.float32)
    dts = np.array([0.005, 0.01, -0.005, 0.0, -0.0, sub, -sub, 1e30, np.inf, -np.inf, np.nan], np.float32)
    assert wlib.wstep_mismatches(ws.ctypes.data, len(ws), dts.ctypes.data, len(dts)) == 0
