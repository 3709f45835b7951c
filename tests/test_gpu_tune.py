"""Launch constants of the hybrid engine (pa_gpu_tracer_set_launch) change
the speed of the fused round, never its results; the round autotuner
(pars_amd.tune.tune_round) times candidates on the caller's round and
leaves the fastest on the tracer."""
import numpy as np
import pytest

import pars_amd as pa
from pars_amd import tune as T

from .helpers import ShareFields, f32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [pa.FP32, pa.BF16])
def test_launch_constants_leave_results_unchanged(ctx, orc, storage):
    S = ShareFields(ctx, orc, 64, 64, 21, storage, layout=pa.LAYOUT_CELLS)
    loc = S.share.local
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    R = 220
    region = R + 40
    seeds = pa.seed_generator.generate(ctx, loc.offset, loc.size, loc.spacing * f32(2), R, 21)
    outs = []
    for q, p, x in [(16, 32, 1), (8, 0, 1), (12, 64, 0), (1, 48, 1), (16, 32, 4), (16, 48, 64)]:
        t.set_launch(q, p)
        t.set_xcd_remap(x)
        r = t.trace_union_round(seeds, region, 2, 0.6,
                                colors=ctx.upload(np.zeros(4 * seeds.n * region, np.float32)))
        nv, ni = r["n_vertices"], r["n_indices"]
        outs.append((r["total_steps"], r["vertices"].download(np.float32, 4 * nv).tobytes(),
                     r["indices"].download(np.int32, ni).tobytes(), r["colors"].download(np.float32, 4 * nv).tobytes(),
                     r["counts"].download(np.int32, seeds.n).tobytes()))
    assert all(o == outs[0] for o in outs[1:])
    with pytest.raises(pa.PaError):
        t.set_launch(17, 32)
    with pytest.raises(pa.PaError):
        t.set_launch(16, -1)


def test_round_autotuner(ctx, orc):
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    loc = S.share.local
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    R = 160
    region = R + 40
    seeds = pa.seed_generator.generate(ctx, loc.offset, loc.size, loc.spacing * f32(2), R, 21)
    b = t.round_buffers(seeds.n, region)
    tun = T.tune_round(ctx, t, seeds, region, b, ctx.alloc(48), quad_switch=(8, 16), pair_live=(32, 64),
                       xcd_remap=(1,), reps=2)
    assert len(tun.table) == 4 and tun.ms == min(tun.table.values())
    assert (tun.quad_switch, tun.pair_live) in [(8, 32), (8, 64), (16, 32), (16, 64)]
    # the tracer keeps the choice, and the round it runs equals the default's
    r = t.trace_union_round(seeds, region, 2, 0.6)
    t.set_launch(16, 32)
    r0 = t.trace_union_round(seeds, region, 2, 0.6)
    assert r["n_vertices"] == r0["n_vertices"] and r["total_steps"] == r0["total_steps"]
    nv = r["n_vertices"]
    assert r["vertices"].download(np.float32, 4 * nv).tobytes() == r0["vertices"].download(np.float32, 4 * nv).tobytes()
