"""The block loader into device bricks (pa_gpu_field_load; data_io.cpp:91-115)
on the committed HDF5 fixture: the decoded nodes are the hyperslab rounded
to the storage, the block's geometry is the host read's, and a trace over
the loaded block equals the oracle's on the same values."""
import os

import numpy as np
import pytest

import pars_amd as pa
from pars_amd import workloads as W

from .golden import make_h5

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors_small.h5")


@pytest.mark.parametrize("storage", [pa.FP32, pa.BF16, pa.FP16])
def test_field_load(ctx, orc, storage):
    offset, ghosted, block = (8, 4, 7), (12, 14, 9), (11, 13, 8)
    f = pa.VectorField.load(ctx, PATH, offset, ghosted, block, storage)
    xyz, off, size, sp = pa.read_vector_block(PATH, offset, ghosted, block)
    assert f.dims == tuple(ghosted)
    assert f.offset.tobytes() == off.tobytes() and f.size.tobytes() == size.tobytes()
    assert f.spacing.tobytes() == sp.tobytes()
    got = f.read()
    assert got.view(np.uint32).tobytes() == W.quantize(xyz, storage).view(np.uint32).tobytes()
    # a round over the loaded block: the oracle on the decoded values
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(f)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.05)
    seeds = orc.seed_generate(off, size, sp * np.float32(2), 40, 3)
    parts = pa.Particles.from_host(ctx, seeds)
    info = t.compute_round_info(parts)
    chunks = t.allocate(info, fill_invalid=True)
    res = t.trace(parts, chunks, info)
    O = orc.Field(got, off, size, sp)
    _, oc, ocnt, _, osteps = orc.trace(seeds, O, dt=0.05, integrator=orc.RK4)
    assert res["total_steps"] == osteps > 0
    assert chunks[0][0].download(np.float32, chunks[0][1] * 4).tobytes() == oc[0].tobytes()
    f.release()
