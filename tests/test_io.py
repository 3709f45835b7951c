"""The block loader's host read (data_io::load_vector_field,
pa/source/stages/data_io.cpp:91-115) on the committed HDF5 fixture
(tests/golden/vectors_small.h5, written by tests/golden/make_h5.py): the
ghosted hyperslab at a rank's offset, the spacing attribute divided by its
largest component, and offset / size in FP32 as data_io.cpp:113-114 has
them.  No GPU: pa_gpu_read_vector_block only reads the file."""
import os

import numpy as np
import pytest

import pars_amd as pa

from .golden import make_h5

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors_small.h5")
f32 = np.float32


@pytest.mark.parametrize("offset,ghosted,block", [((0, 0, 0), (20, 18, 16), (20, 18, 16)),
                                                  ((0, 0, 0), (9, 10, 8), (8, 9, 8)),
                                                  ((8, 4, 7), (12, 14, 9), (11, 13, 8)),
                                                  ((19, 17, 15), (1 + 1, 1 + 1, 1 + 1), (1, 1, 1))])
def test_read_hyperslab(offset, ghosted, block):
    if offset == (19, 17, 15):  # a 2^3 slab would reach past the data: rejected
        with pytest.raises(pa.PaError) as e:
            pa.read_vector_block(PATH, offset, ghosted, block)
        assert e.value.status == pa.E_INVALID and "outside" in str(e.value)
        return
    xyz, off, size, sp = pa.read_vector_block(PATH, offset, ghosted, block)
    data = make_h5.vectors_data()
    o, g = offset, ghosted
    want = data[o[0]:o[0] + g[0], o[1]:o[1] + g[1], o[2]:o[2] + g[2]]
    assert xyz.tobytes() == np.ascontiguousarray(want).tobytes()
    want_sp = make_h5.SPACING / make_h5.SPACING.max()  # FP32 division by maxCoeff
    assert sp.tobytes() == want_sp.astype(np.float32).tobytes()
    assert off.tobytes() == (np.array(offset, np.int32).astype(np.float32) * want_sp).astype(np.float32).tobytes()
    assert size.tobytes() == (np.array(block, np.int32).astype(np.float32) * want_sp).astype(np.float32).tobytes()


def test_read_errors(tmp_path):
    with pytest.raises(pa.PaError) as e:
        pa.read_vector_block(str(tmp_path / "missing.h5"), (0, 0, 0), (4, 4, 4), (3, 3, 3))
    assert "cannot open" in str(e.value)
    with pytest.raises(pa.PaError):
        pa.read_vector_block(PATH, (-1, 0, 0), (4, 4, 4), (3, 3, 3))
    with pytest.raises(pa.PaError):
        pa.read_vector_block(PATH, (0, 0, 0), (1, 4, 4), (1, 3, 3))


def test_fixture_is_what_the_script_writes(tmp_path):
    """The committed file equals a fresh write of the script (so the data the
    tests expect is the data in the file)."""
    fresh = str(tmp_path / "v.h5")
    make_h5.write(fresh)
    a, *_ = pa.read_vector_block(PATH, (0, 0, 0), (20, 18, 16), (20, 18, 16))
    b, *_ = pa.read_vector_block(fresh, (0, 0, 0), (20, 18, 16), (20, 18, 16))
    assert a.tobytes() == b.tobytes()
