"""pa_gpu_finish_counted (counted prune, index generation and colouring of
a chunk in one pass) against the three stages it fuses, run one after the
other on the same slots: pa_gpu_prune_counted, pa_gpu_index_generate
(index_generator.cpp:10-24) and pa_gpu_color_generate
(color_generator.cpp:16-98).  Vertices and indices must be equal byte for
byte, and colours byte for byte at every position the colour stage writes.
Chunks with curves of the shape the tracers write take the fused pass; the
irregular chunks (a seed alone, a termination-valued vertex inside a curve,
a curve without its termination vertex) take the general kernels."""
import ctypes as C

import numpy as np
import pytest

import pars_amd as pa
from pars_amd import lib

pytestmark = pytest.mark.gpu

TERM = np.array([-2.0, -2.0, -2.0, 0.0], np.float32)


def _slots(rng, n, R, irregular=None):
    counts = rng.integers(2, R + 1, n).astype(np.int32)
    counts[rng.random(n) < 0.2] = 2  # seeds that end at once (seed + termination vertex)
    slots = rng.standard_normal((n, R, 4)).astype(np.float32)
    slots[:, :, 3] = 0.0
    for p in range(n):
        slots[p, counts[p] - 1] = TERM
    if irregular == "alone":
        counts[n // 2] = 1
    elif irregular == "inside":
        p = int(np.argmax(counts))
        slots[p, counts[p] // 2] = TERM
    elif irregular == "open":
        p = int(np.argmax(counts))
        slots[p, counts[p] - 1] = [0.5, 0.25, 0.125, 0.0]
    elif irregular == "last_final":
        counts[-5:] = 2  # the chunk's final index is not in its last seed
    return slots, counts


def _separate(ctx, d_slots, R, d_counts, n, cap, mode, param):
    v = ctx.alloc(cap * 16)
    idx = ctx.alloc(cap * 4)
    col = ctx.upload(np.full((cap, 4), np.nan, np.float32))
    m, ni = C.c_size_t(), C.c_size_t()
    ctx.check(lib().pa_gpu_prune_counted(ctx.handle, C.c_void_p(d_slots.ptr), C.c_size_t(R),
                                         C.c_void_p(d_counts.ptr), C.c_size_t(n), C.c_void_p(v.ptr),
                                         C.c_size_t(cap), C.byref(m), None), "prune_counted")
    ctx.check(lib().pa_gpu_index_generate(ctx.handle, C.c_void_p(v.ptr), m, C.c_void_p(idx.ptr), C.byref(ni)),
              "index")
    ctx.check(lib().pa_gpu_color_generate(ctx.handle, C.c_void_p(v.ptr), m, C.c_void_p(idx.ptr), ni, C.c_int(mode),
                                          C.c_float(param), C.c_void_p(col.ptr)), "color")
    return (v.download(np.float32, 4 * m.value), idx.download(np.int32, ni.value),
            col.download(np.float32, 4 * m.value).reshape(-1, 4))


def _fused(ctx, d_slots, R, d_counts, n, cap, mode, param):
    v = ctx.alloc(cap * 16)
    idx = ctx.alloc(cap * 4)
    col = ctx.upload(np.full((cap, 4), 7.0, np.float32))
    offs = ctx.alloc((n + 1) * 8)
    m, ni = C.c_size_t(), C.c_size_t()
    ctx.check(lib().pa_gpu_finish_counted(ctx.handle, C.c_void_p(d_slots.ptr), C.c_size_t(R),
                                          C.c_void_p(d_counts.ptr), C.c_size_t(n), C.c_void_p(v.ptr),
                                          C.c_size_t(cap), C.byref(m), C.c_void_p(offs.ptr), C.c_void_p(idx.ptr),
                                          C.byref(ni), C.c_int(mode), C.c_float(param), C.c_void_p(col.ptr)),
              "finish_counted")
    return (v.download(np.float32, 4 * m.value), idx.download(np.int32, ni.value),
            col.download(np.float32, 4 * m.value).reshape(-1, 4), offs.download(np.int64, n + 1))


@pytest.mark.parametrize("irregular", [None, "last_final", "alone", "inside", "open"])
@pytest.mark.parametrize("mode,param", [(pa.COLOR_HSV_CONSTANT_S, 0.75), (pa.COLOR_RGB, 0.0),
                                        (pa.COLOR_MAGNITUDE, 2.5)])
@pytest.mark.parametrize("n,R", [(1, 3), (37, 5), (3000, 150)])
def test_finish_counted_equals_stages(ctx, n, R, mode, param, irregular):
    rng = np.random.default_rng(n * 1000 + R)
    slots, counts = _slots(rng, n, R, irregular)
    d_slots, d_counts = ctx.upload(slots), ctx.upload(counts)
    cap = int(counts.sum()) + 8
    v0, i0, c0 = _separate(ctx, d_slots, R, d_counts, n, cap, mode, param)
    v1, i1, c1, offs = _fused(ctx, d_slots, R, d_counts, n, cap, mode, param)
    assert v1.tobytes() == v0.tobytes()
    assert i1.tobytes() == i0.tobytes()
    assert (offs == np.concatenate([[0], np.cumsum(counts.astype(np.int64))])).all()
    written = ~np.isnan(c0).any(axis=1)
    assert c1[written].tobytes() == c0[written].tobytes()
    if irregular is None or irregular == "last_final":
        # the fused pass writes exactly the stage's positions: the others keep what the buffer held
        assert (c1[~written] == 7.0).all()


def test_finish_counted_errors(ctx):
    counts = np.array([3, 4], np.int32)
    slots = np.zeros((2, 4, 4), np.float32)
    d_slots, d_counts = ctx.upload(slots), ctx.upload(counts)
    buf = ctx.alloc(64 * 16)
    m, ni = C.c_size_t(), C.c_size_t()
    st = lib().pa_gpu_finish_counted(ctx.handle, C.c_void_p(d_slots.ptr), C.c_size_t(4), C.c_void_p(d_counts.ptr),
                                     C.c_size_t(2), C.c_void_p(buf.ptr), C.c_size_t(6), C.byref(m), None,
                                     C.c_void_p(buf.ptr), C.byref(ni), C.c_int(pa.COLOR_RGB), C.c_float(0.0),
                                     C.c_void_p(buf.ptr))
    assert st == pa.E_CAPACITY
    st = lib().pa_gpu_finish_counted(ctx.handle, C.c_void_p(d_slots.ptr), C.c_size_t(4), C.c_void_p(d_counts.ptr),
                                     C.c_size_t(2), C.c_void_p(buf.ptr), C.c_size_t(64), C.byref(m), None,
                                     C.c_void_p(buf.ptr), C.byref(ni), C.c_int(77), C.c_float(0.0),
                                     C.c_void_p(buf.ptr))
    assert st == pa.E_INVALID
