"""Batched seed sets with their curves (pa_gpu_pool_run_vertices and the
streamed driver pars_amd.batch.run_batch_vertices): every set of a small
uneven batch, finished as soon as its particles have all ended, in chunks
as the reference chunks its curves, equals the oracle's one-block round of
that set alone (particle_tracer.cpp:243-337: slots, prune :402-409),
index_generator and color_generator on each chunk (vertices and indices
bit for bit, colours within 1e-6 as the colour stage's own parity test).  The arena
holds only part of the batch, so sets are admitted as others finish, and
passes are bounded, so particles resume where they left off."""
import numpy as np
import pytest

import pars_amd as pa
from pars_amd import batch as B

pytestmark = pytest.mark.gpu


def _written(idx, n):
    w = np.zeros(n, bool)
    if len(idx):
        w[idx] = True
        ends = np.nonzero(np.append(idx[1:] != idx[:-1] + 1, True))[0][:-1]
        w[idx[ends] + 1] = True
    return w


@pytest.mark.parametrize("storage,layout", [(pa.BF16, pa.LAYOUT_CELLS), (pa.FP32, pa.LAYOUT_NODES),
                                            (pa.FP32, pa.LAYOUT_CELLS)])
def test_batch_vertices_equal_reference_rounds(ctx, orc, storage, layout):
    wl = B.make_workload(grids=(16, 24, 20), n_sets=6, n_modes=8, size_range=(60, 700), steps_range=(20, 150))
    fields = B.build_fields(ctx, wl, storage, layout)
    ofields = [orc.Field(f.read(), f.offset, f.size, f.spacing) for f in fields]
    sizes = np.array([s["size"] for s in wl.sets])
    Rs = np.array([s["steps"] + 2 for s in wl.sets])
    need = sizes * Rs
    got = {}

    def consume(c):
        v = ctx_download(c.vertices, np.float32, 4 * c.n_vertices).reshape(-1, 4)
        i = ctx_download(c.indices, np.int32, c.n_indices)
        col = ctx_download(c.colors, np.float32, 4 * c.n_vertices).reshape(-1, 4)
        got.setdefault(c.set_id, []).append((c.chunk, c.seed_begin, c.seed_end, v, i, col))

    def ctx_download(ptr, dtype, count):
        out = np.empty(count, dtype)
        if count:
            ctx.check(pa.lib().pa_gpu_memcpy_d2h(ctx.handle, out.ctypes.data_as(pa.C.c_void_p), pa.C.c_void_p(ptr),
                                                 pa.C.c_size_t(out.nbytes)), "d2h")
        return out

    chunk = 40 * int(Rs.max())  # several chunks per set
    tot = B.run_batch_vertices(ctx, wl, fields, waves_per_cu=8, arena_slots=int(need.max() * 1.3),
                               pass_iterations=37, consumer=consume, chunk_vertices=chunk)
    assert tot["sets"] == len(sizes) and tot["passes"] > 1
    order = B.set_order(wl)
    parts = wl.particles[order]
    off = np.concatenate([[0], np.cumsum(sizes)])
    steps_all = 0
    for s in range(len(sizes)):
        p = parts[off[s]:off[s + 1]].copy()
        fid = int(p["vector_field_index"][0])
        p["vector_field_index"] = -1  # the oracle's own block
        info, chunks, counts, hand, steps = orc.trace(p, ofields[fid], dt=B.DT, integrator=orc.RK4)
        steps_all += steps
        slots = np.concatenate(chunks).reshape(len(p), -1, 4)[:, :Rs[s]]
        per = chunk // int(Rs[s])
        parts_got = sorted(got[s])
        assert [c[1] for c in parts_got] == list(range(0, len(p), per))
        for k, j0, j1, v, i, col in parts_got:
            want_v = orc.prune(slots[j0:j1].reshape(-1, 4))
            assert v.tobytes() == want_v.tobytes(), (s, k)
            want_i = orc.index_generate(want_v) if len(want_v) else np.zeros(0, np.int32)
            assert i.tobytes() == want_i.tobytes(), (s, k)
            if len(want_i):
                want_c = orc.color_generate(want_v, want_i, mode=2, parameter=0.75)
                w = _written(want_i, len(want_v))
                # the colour stage against the oracle: within 1e-6 (HSV; the trigonometry differs in the
                # last bits), as test_prune_index_color_parity has it
                np.testing.assert_allclose(col[w], want_c[w], atol=1e-6, rtol=0, err_msg=str((s, k)))
    assert tot["steps"] == steps_all
    for f in fields:
        f.release()
