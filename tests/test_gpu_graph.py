"""Graph-captured rounds: the seed lattice plus the asynchronous fused round,
and the batched pool's initialisation plus run, captured once and replayed.
Every replay, with the program's own eager work and synchronisations between
replays, equals the eager round bit for bit (vertices, indices, colours,
counts, result words; pool endpoints, steps and request counters)."""
import numpy as np
import pytest

import pars_amd as pa

from .helpers import ShareFields, f32

pytestmark = pytest.mark.gpu


def test_captured_round_replays(ctx, orc):
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    sh = S.share
    R, dt = 200, 0.02
    t = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    loc = sh.local
    stride = loc.spacing * f32(2)
    n, _ = pa.seed_generator.count(loc.size, stride)
    region = R + 40
    # eager reference
    seeds = pa.seed_generator.generate(ctx, loc.offset, loc.size, stride, R, 21)
    want = t.trace_union_round(seeds, region, 2, 0.6, colors=ctx.upload(np.zeros(4 * n * region, np.float32)))
    nv, ni = want["n_vertices"], want["n_indices"]
    want_v = want["vertices"].download(np.float32, 4 * nv)
    want_i = want["indices"].download(np.int32, ni)
    want_c = want["colors"].download(np.float32, 4 * nv)
    want_cnt = want["counts"].download(np.int32, n)
    # captured: seeds + asynchronous round into their own buffers
    sbuf = ctx.upload(seeds.download())
    b = t.round_buffers(n, region)
    result = ctx.alloc(48)
    parts = pa.Particles(sbuf, n)
    t.trace_union_round_async(parts, region, b, result, 2, 0.6)  # sizes the workspace before the capture
    # colours the rule leaves alone keep what the buffer held: start from zeros, as the eager round did
    b["out_colors"].upload(np.zeros(4 * n * region, np.float32))
    sbuf.upload(np.zeros(n, pa.PARTICLE_DTYPE))  # the captured seed kernel refills it
    ctx.synchronize()
    lib, C = pa.lib(), pa.C
    with pa.Graph.capture(ctx) as g:
        m = C.c_size_t()
        ctx.check(lib.pa_gpu_seed_generate(ctx.handle, pa._f3(loc.offset), pa._f3(loc.size), pa._f3(stride),
                                           C.c_int32(R), C.c_int32(21), C.c_void_p(sbuf.ptr), C.c_size_t(n),
                                           C.byref(m)), "seeds")
        t.trace_union_round_async(parts, region, b, result, 2, 0.6)
    for k in range(4):
        # eager work of the program's own between replays: scribble on the outputs
        b["out_vertices"].upload(np.full(4 * 64, 9.0, np.float32))
        b["out_indices"].upload(np.full(64, -7, np.int32))
        result.upload(np.zeros(6, np.uint64))
        ctx.synchronize()
        g.launch()
        ctx.synchronize()
        rr = result.download(pa.ROUND_RESULT_DTYPE, 1)
        r = rr[0]
        assert (r["n_vertices"], r["n_indices"], r["total_steps"]) == (nv, ni, want["total_steps"]), k
        assert r["n_continuation"] == 0 and r["flags"] == 0
        ctx.check(lib.pa_gpu_round_result_check(ctx.handle, rr.ctypes.data_as(C.c_void_p), C.c_size_t(n * region),
                                                C.c_size_t(n * region)), "result")
        assert b["out_vertices"].download(np.float32, 4 * nv).tobytes() == want_v.tobytes(), k
        assert b["out_indices"].download(np.int32, ni).tobytes() == want_i.tobytes(), k
        assert b["out_colors"].download(np.float32, 4 * nv).tobytes() == want_c.tobytes(), k
        assert np.array_equal(b["counts"].download(np.int32, n), want_cnt), k
    g.close()


def test_capture_rejects_read_back(ctx, orc):
    """A call that reads back inside a capture fails the capture; the context
    works normally afterwards."""
    d = pa.VectorField.from_host(ctx, np.ones((8, 8, 8, 3), np.float32), (0, 0, 0), (8, 8, 8), (1, 1, 1))
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.1)
    seeds = pa.Particles.from_host(ctx, orc.seed_generate((0, 0, 0), (8, 8, 8), (2, 2, 2), 10, 0))
    with pytest.raises(pa.PaError):
        with pa.Graph.capture(ctx):
            t.trace_union_round(seeds, 16)  # synchronous: reads back
    r = t.trace_union_round(seeds, 16)
    assert r["n_vertices"] > 0


def test_captured_batch_replays(ctx, orc):
    """The batched pool (several seed sets over two fields) as a captured
    round: init (clearing the request counters) + one persistent run."""
    from .test_gpu_parity import _pool_fields, _pool_particles
    fields, fo = _pool_fields(ctx, orc, pa.BF16, pa.LAYOUT_CELLS)
    parts = _pool_particles(orc, fo, 900, 5)
    parts["original_rank"] = np.arange(len(parts)) % 3  # three seed sets
    n = len(parts)
    t = pa.particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    dev = pa.Particles.from_host(ctx, parts)
    eager = pa.ParticlePool(ctx, n, requests=3)
    eager.init(dev, 2)
    steps = eager.run(t, fields)
    want = eager.results()
    want_done = eager.request_done.download(np.int32, 3)
    pool = pa.ParticlePool(ctx, n, requests=3)
    rej = ctx.alloc(8)
    res = ctx.alloc(16)
    with pa.Graph.capture(ctx) as g:
        pa.pool_init_async(ctx, pool, dev, 2, rej, zero_requests=True)
        pa.pool_run_async(pool, t, fields, res, n=n)
    for k in range(3):
        pool.position.upload(np.zeros(4 * 16, np.float32))
        pool.request_done.upload(np.full(3, 99, np.int32))
        ctx.synchronize()
        g.launch()
        ctx.synchronize()
        got = pool.results(n)
        assert got["position"].tobytes() == want["position"].tobytes(), k
        assert np.array_equal(got["steps"], want["steps"]) and (got["status"] == pa.POOL_DONE).all()
        assert np.array_equal(pool.request_done.download(np.int32, 3), want_done)
        assert res.download(np.uint64, 2).tolist() == [steps, 0]
        assert rej.download(np.uint32, 1)[0] == 0
    g.close()
