"""Parity of the headline round at its full size: one rank's share of a 512^3
ABC domain split 4x4x4 (rank 21: its own 129^3 ghosted block and six face
neighbours), 262,144 seeds on the stride-2 lattice, runge_kutta_4 with
dt 0.005 and remaining_iterations 1002, stop rule |u| < 1e-6, union-of-7
continuation; the fused round (trace + prune + index + colour) through the
C ABI against the oracle's union trace, index generator and colour
generator on the same decoded node values.

Vertices, per-seed slot counts, step counts and indices are compared byte
for byte (particle_tracer.cpp:253-337 followed by prune :402-409,
index_generator.cpp:10-24); colours (color_generator.cpp:16-40,
hsv_constant_s 0.75) at 1e-6 absolute, since device and glibc atan2f/acosf
differ in the last bits.  Runs for FP32 cell records (the bench's form),
FP32 nodes and BF16/FP16 cell records (the oracle reads the 16-bit values
as decoded to FP32)."""
import os

import numpy as np
import pytest

import pars_amd as pa
from pars_amd import workloads as W

pytestmark = pytest.mark.gpu

R = 1002          # 1000 integrator steps
DT = 0.005
EXTRA_SLOTS = 30  # per-seed room for the union's extra segment vertices
MODE, PARAM = 2, 0.75  # hsv_constant_s


@pytest.fixture(scope="module")
def share():
    return W.share(512, 64, 21)


@pytest.mark.parametrize("storage,layout", [(pa.FP32, pa.LAYOUT_CELLS), (pa.FP32, pa.LAYOUT_NODES),
                                            (pa.BF16, pa.LAYOUT_CELLS), (pa.FP16, pa.LAYOUT_CELLS)],
                         ids=["fp32-cells", "fp32-nodes", "bf16-cells", "fp16-cells"])
def test_headline_round_fullsize(ctx, orc, share, storage, layout):
    orc.set_threads(min(16, os.cpu_count() or 1))
    blocks = [share.local] + list(share.neighbors)
    fields = [pa.VectorField.synthesize(ctx, "abc", W.ABC, b.ghosted, b.int_offset, share.h, b.offset, b.size,
                                        b.spacing, storage).set_layout(layout) for b in blocks]
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(fields[0])
    t.set_neighbor_vector_fields(fields[1:], share.neighbor_ranks)
    t.set_neighbor_face_masks(share.face_masks)
    t.set_integrator(pa.RUNGE_KUTTA_4)
    t.set_step_size(DT)
    t.set_stop_rule(pa.STOP_NORM, 1e-6)
    t.set_engine(pa.ENGINE_HYBRID)
    loc = share.local
    seeds = pa.seed_generator.generate(ctx, loc.offset, loc.size, loc.spacing * np.float32(2), R, 21)
    n = seeds.n
    assert n == 262144
    region = R + EXTRA_SLOTS
    r = t.trace_union_round(seeds, region, MODE, PARAM)
    assert r["continuation"].n == 0
    nv, ni = r["n_vertices"], r["n_indices"]
    got_v = r["vertices"].download(np.float32, 4 * nv).reshape(nv, 4)
    got_cnt = r["counts"].download(np.int32, n)
    got_steps = r["seed_steps"].download(np.int32, n)
    got_i = r["indices"].download(np.int32, ni)
    got_c = r["colors"].download(np.float32, 4 * nv).reshape(nv, 4)
    seeds_np = seeds.download()
    ofields = [orc.Field(f.read(), b.offset, b.size, b.spacing) for f, b in zip(fields, blocks)]
    for f in fields:
        f.release()

    ou = orc.trace_union(seeds_np, ofields[0], ofields[1:], share.neighbor_ranks, share.face_masks, dt=DT,
                         integrator=orc.RK4, stop_rule=orc.STOP_NORM, stop_eps=1e-6, vertex_capacity=n * region)
    want_v = ou["vertices"]
    assert r["total_steps"] == ou["total_steps"] > 40_000_000
    assert int(got_steps.astype(np.int64).sum()) == ou["total_steps"]
    want_cnt = np.bincount(ou["segment_seed"], weights=ou["segment_length"], minlength=n).astype(np.int64)
    assert np.array_equal(got_cnt.astype(np.int64), want_cnt)
    assert nv == len(want_v)
    if got_v.tobytes() != want_v.tobytes():
        bad = np.nonzero(np.any(got_v.view(np.uint32) != want_v.view(np.uint32), axis=1))[0]
        seed_of = np.repeat(np.arange(n), want_cnt)
        pytest.fail("%d of %d vertices differ (first at %d, seed %d)" % (len(bad), nv, bad[0], seed_of[bad[0]]))
    assert (ou["segment_block"] >= 0).any()  # some seeds crossed into neighbour blocks
    want_i = orc.index_generate(want_v)
    assert ni == len(want_i)
    assert np.array_equal(got_i, want_i)
    want_c = orc.color_generate(want_v, want_i, MODE, PARAM)
    # only the colours the reference writes: each index, and the vertex after
    # an index that is not followed by the next index (the last one excepted)
    written = np.zeros(nv, bool)
    written[want_i] = True
    ends = want_i[:-1][want_i[1:] != want_i[:-1] + 1]
    written[ends + 1] = True
    diff = np.abs(got_c[written].astype(np.float64) - want_c[written].astype(np.float64))
    assert float(diff.max()) <= 1e-6, float(diff.max())
