"""The fused round without allocation (pa_gpu_tracer_trace_union_round_noalloc)
equals the fused round (pa_gpu_tracer_trace_union_round) bit for bit:
slots, counts, step count, vertices, indices, colours, endpoints and seed
steps, on two calls in a row, the second with fewer seeds (the fixed
workspace keeps the first call's contents), and it rejects a call whose
workspace would not fit without writing anything."""
import numpy as np
import pytest

import pars_amd as pa

from .helpers import ShareFields, f32

pytestmark = pytest.mark.gpu

C = pa.C


def _round(ctx, t, fn, parts, n, region, b, mode, param):
    ncont, nv, ni, steps = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
    st = fn(t._h, C.c_void_p(parts.ptr), C.c_size_t(n), C.c_size_t(region), C.c_void_p(b["vertices"].ptr),
            C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.byref(ncont), C.c_int(mode),
            C.c_float(param), C.c_void_p(b["out_vertices"].ptr), C.c_size_t(b["cap"]), C.byref(nv),
            C.c_void_p(b["out_indices"].ptr), C.c_size_t(b["cap"]), C.byref(ni), C.c_void_p(b["out_colors"].ptr),
            C.c_void_p(b["endpoints"].ptr), C.c_void_p(b["seed_steps"].ptr), C.byref(steps))
    return st, dict(n_cont=ncont.value, nv=nv.value, ni=ni.value, steps=steps.value)


def _buffers(ctx, t, n, region, fill):
    b = t.union_buffers(n, region, endpoints=True)
    b["cap"] = n * region
    b["out_vertices"] = ctx.upload(np.full(4 * b["cap"], fill, np.float32))
    b["out_indices"] = ctx.alloc(4 * b["cap"])
    b["out_colors"] = ctx.upload(np.full(4 * b["cap"], fill, np.float32))
    return b


def _outputs(b, n, r):
    n8 = (n + 7) // 8 * 8
    return dict(counts=b["counts"].download(np.int32, n).tobytes(),
                endpoints=b["endpoints"].download(np.float32, 4 * n).tobytes(),
                seed_steps=b["seed_steps"].download(np.int32, n).tobytes(),
                vertices=b["out_vertices"].download(np.float32, 4 * r["nv"]).tobytes(),
                indices=b["out_indices"].download(np.int32, r["ni"]).tobytes(),
                colors=b["out_colors"].download(np.float32, 4 * r["nv"]).tobytes(),
                slots_len=n8)


def test_noalloc_round_equals_round(ctx, orc):
    lib = pa.lib()
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    sh = S.share
    dt, R = 0.02, 200
    t = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing, R, 21)  # 16^3 seeds
    seeds["remaining_iterations"][::97] = 1
    region = R + 40
    for n in (len(seeds), 1001):  # the second call is smaller
        print("noalloc round: n =", n, flush=True)
        parts = pa.Particles.from_host(ctx, seeds[:n])
        want_b = _buffers(ctx, t, n, region, 7.5)
        got_b = _buffers(ctx, t, n, region, 7.5)
        st, want = _round(ctx, t, lib.pa_gpu_tracer_trace_union_round, parts, n, region, want_b, 2, 0.75)
        assert st == pa.OK
        st, got = _round(ctx, t, lib.pa_gpu_tracer_trace_union_round_noalloc, parts, n, region, got_b, 2, 0.75)
        assert st == pa.OK, pa.lib().pa_gpu_last_error(ctx.handle)
        assert got == want and want["nv"] > n and want["steps"] > 0
        assert _outputs(got_b, n, got) == _outputs(want_b, n, want)


def test_noalloc_round_rejects_oversize(ctx, orc):
    """A call whose workspace exceeds the fixed array fails with
    PA_E_CAPACITY before any kernel runs: the outputs keep their contents."""
    lib = pa.lib()
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    n = 1 << 20  # ~29 MiB of workspace
    sh = S.share
    one = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(8), 4, 21)
    seeds = np.resize(one, n)
    parts = pa.Particles.from_host(ctx, seeds)
    region = 4
    b = _buffers(ctx, t, n, region, 3.25)
    st, r = _round(ctx, t, lib.pa_gpu_tracer_trace_union_round_noalloc, parts, n, region, b, 2, 0.75)
    assert st == pa.E_CAPACITY
    assert r["nv"] == 0 and r["ni"] == 0
    msg = lib.pa_gpu_last_error(ctx.handle)
    assert b"seeds per call" in msg
    assert (b["out_vertices"].download(np.float32, 4 * 64) == 3.25).all()
    # and the workspace-sized call after it works
    m = 5000
    want_b, got_b = _buffers(ctx, t, m, region, 0.0), _buffers(ctx, t, m, region, 0.0)
    p2 = pa.Particles.from_host(ctx, seeds[:m])
    st, want = _round(ctx, t, lib.pa_gpu_tracer_trace_union_round, p2, m, region, want_b, 2, 0.75)
    st2, got = _round(ctx, t, lib.pa_gpu_tracer_trace_union_round_noalloc, p2, m, region, got_b, 2, 0.75)
    assert st == st2 == pa.OK and got == want
    assert _outputs(got_b, m, got) == _outputs(want_b, m, want)
