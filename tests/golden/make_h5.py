"""Writes tests/golden/vectors_small.h5, the block loader's fixture, with
libhdf5 through ctypes (h5py is not available): the dataset "vectors"
[20][18][16][3] float32 (seeded normal values, see vectors_data()) and the
file attribute "spacing" = (0.5, 0.25, 0.75), the layout pars' data_io reads
(pa/source/stages/data_io.cpp:91-115).  Run from the repository root:
python tests/golden/make_h5.py"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "vectors_small.h5")
SHAPE = (20, 18, 16, 3)
SPACING = np.array([0.5, 0.25, 0.75], np.float32)


def vectors_data():
    return np.random.default_rng(20261204).standard_normal(SHAPE).astype(np.float32)


def _hdf5():
    for name in ("libhdf5.so", "/opt/conda/lib/libhdf5.so"):
        try:
            return C.CDLL(name)
        except OSError:
            pass
    raise OSError("libhdf5 not found")


def write(path=PATH):
    h = _hdf5()
    hid = C.c_int64
    for f in ("H5Fcreate", "H5Screate_simple", "H5Dcreate2", "H5Acreate2"):
        getattr(h, f).restype = hid
    assert h.H5open() >= 0
    f32_le = hid.in_dll(h, "H5T_IEEE_F32LE_g")
    native = hid.in_dll(h, "H5T_NATIVE_FLOAT_g")
    data = vectors_data()
    f = h.H5Fcreate(path.encode(), C.c_uint(2), hid(0), hid(0))  # H5F_ACC_TRUNC, default property lists
    assert f >= 0
    dims = (C.c_ulonglong * 4)(*SHAPE)
    space = h.H5Screate_simple(4, dims, None)
    dset = h.H5Dcreate2(hid(f), b"vectors", f32_le, hid(space), hid(0), hid(0), hid(0))
    assert dset >= 0
    assert h.H5Dwrite(hid(dset), native, hid(0), hid(0), hid(0), data.ctypes.data_as(C.c_void_p)) >= 0
    adims = (C.c_ulonglong * 1)(3)
    aspace = h.H5Screate_simple(1, adims, None)
    attr = h.H5Acreate2(hid(f), b"spacing", f32_le, hid(aspace), hid(0), hid(0))
    assert attr >= 0
    assert h.H5Awrite(hid(attr), native, SPACING.ctypes.data_as(C.c_void_p)) >= 0
    for fn, x in (("H5Aclose", attr), ("H5Sclose", aspace), ("H5Dclose", dset), ("H5Sclose", space), ("H5Fclose", f)):
        assert getattr(h, fn)(hid(x)) >= 0


if __name__ == "__main__":
    write()
    print("wrote", PATH)
