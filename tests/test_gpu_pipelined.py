"""Pipelined rounds (pa_gpu_tracer_trace_union_round_pipelined): the finish
of a round runs on the context's second stream while the next round traces.
Rounds of different seed sets queued back to back, each on its own buffers
or on two alternated sets, equal the synchronous fused round bit for bit
(vertices, indices, colours, counts and the result words), and later work on
the context is ordered after their finishes by pa_gpu_round_join."""
import numpy as np
import pytest

import pars_amd as pa

from .helpers import ShareFields, f32

pytestmark = pytest.mark.gpu


def _eager(t, seeds, region):
    n = seeds.n
    want = t.trace_union_round(seeds, region, 2, 0.6, colors=t.ctx.upload(np.zeros(4 * n * region, np.float32)))
    nv, ni = want["n_vertices"], want["n_indices"]
    return dict(v=want["vertices"].download(np.float32, 4 * nv), i=want["indices"].download(np.int32, ni),
                c=want["colors"].download(np.float32, 4 * nv), cnt=want["counts"].download(np.int32, n),
                steps=want["total_steps"], nv=nv, ni=ni)


def _written(idx):
    """Colour positions the colour rule writes (color_generator.cpp:16-98): each index's vertex,
    and the vertex after it where a curve ends (but not after the last index)."""
    w = np.zeros(int(idx.max()) + 2 if len(idx) else 0, bool)
    w[idx] = True
    ends = np.nonzero(idx[1:] != idx[:-1] + 1)[0]
    w[idx[ends] + 1] = True
    return np.nonzero(w)[0]


def _check(ctx, b, res, want, n, region, tag, all_colours):
    r = res.download(pa.ROUND_RESULT_DTYPE, 1)[0]
    assert (r["n_vertices"], r["n_indices"], r["total_steps"]) == (want["nv"], want["ni"], want["steps"]), tag
    assert r["n_continuation"] == 0 and r["flags"] == 0, tag
    nv, ni = want["nv"], want["ni"]
    assert b["out_vertices"].download(np.float32, 4 * nv).tobytes() == want["v"].tobytes(), tag
    assert b["out_indices"].download(np.int32, ni).tobytes() == want["i"].tobytes(), tag
    got_c = b["out_colors"].download(np.float32, 4 * nv).reshape(nv, 4)
    want_c = want["c"].reshape(nv, 4)
    # a reused set keeps an older round's colours where the rule writes none
    sel = slice(None) if all_colours else _written(want["i"])
    assert got_c[sel].tobytes() == want_c[sel].tobytes(), tag
    assert np.array_equal(b["counts"].download(np.int32, n), want["cnt"]), tag


@pytest.mark.parametrize("n_sets", [4, 2])
def test_pipelined_rounds_equal_synchronous(ctx, orc, n_sets):
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    loc = S.share.local
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    Rs = [200, 120, 260, 90, 230]
    region = max(Rs) + 40
    seeds = [pa.seed_generator.generate(ctx, loc.offset, loc.size, loc.spacing * f32(s), R, 21)
             for s, R in zip([2, 3, 2, 4, 3], Rs)]
    want = [_eager(t, sd, region) for sd in seeds]
    nmax = max(sd.n for sd in seeds)
    sets = [t.round_buffers(nmax, region) for _ in range(n_sets)]
    results = [ctx.alloc(48) for _ in range(n_sets)]
    for b in sets:  # colours the rule leaves alone keep what the buffer held: start from zeros
        b["out_colors"].upload(np.zeros(4 * nmax * region, np.float32))
    ctx.synchronize()
    for k, sd in enumerate(seeds):
        b = sets[k % n_sets]
        # with two sets, round k reuses round k-2's buffers: the library orders this trace after that finish
        t.trace_union_round_pipelined(sd, region, b, results[k % n_sets], 2, 0.6, vertex_capacity=nmax * region)
    ctx.round_join()
    # ordered after the finishes: a copy on the context's stream sees the last rounds' outputs
    last = len(seeds) - 1
    r = results[last % n_sets].download(pa.ROUND_RESULT_DTYPE, 1)[0]
    assert r["n_vertices"] == want[last]["nv"]
    ctx.synchronize()
    for k in range(len(seeds)):
        if k + n_sets < len(seeds):
            continue  # overwritten by a later round on the same set
        _check(ctx, sets[k % n_sets], results[k % n_sets], want[k], seeds[k].n, region, k, all_colours=k < n_sets)
