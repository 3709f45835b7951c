"""The tracer's fallbacks to IEEE division, bit for bit against the oracle.

The device takes the corrected quotient (p - o) / h and the corrected weight
fmod(p - o, h) / h everywhere and divides only where some lane of the wave
needs it: a block spacing outside [2^-20, 2^20], a coordinate within 2^-40
of the block's lower face, or a remainder below 2^-40 (a point just off a
node plane).  These cases are rare in the share workload, so they are placed
here on purpose, for the interpolation kernel, the reference round and the
fused round (hybrid engine: lane mode, then quad mode, on cell records)."""
import numpy as np
import pytest

import pars_amd as pa
from tests.helpers import f32, single_block

pytestmark = pytest.mark.gpu

TINY = [f32(2.0 ** -45), f32(-(2.0 ** -45)), np.uint32(7).view(np.float32), f32(2.0 ** -41)]


def near_planes(h, n, count, rng):
    """Points on node planes of an n^3 block, some coordinates nudged by less than 2^-40."""
    k = rng.integers(0, n - 1, size=(count, 3)).astype(np.float32)
    pts = np.zeros((count, 4), np.float32)
    pts[:, :3] = k * h
    nudge = rng.integers(0, len(TINY) + 1, size=(count, 3))
    for j, t in enumerate(TINY):
        pts[:, :3] = np.where(nudge == j, pts[:, :3] + t, pts[:, :3])
    return pts


@pytest.mark.parametrize("h", [None, 2.0 ** -21, 1.5 * 2.0 ** 20], ids=["workload", "below", "above"])
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
def test_interpolate_fallbacks(ctx, orc, h, layout):
    n = 17
    d, o, _ = single_block(ctx, orc, n, h=h, layout=layout)
    hs = o.spacing[0]
    rng = np.random.default_rng(31)
    pts = np.concatenate([near_planes(hs, n, 6000, rng),
                          np.pad(rng.uniform(0, f32(n - 1) * hs, size=(2000, 3)).astype(np.float32), ((0, 0), (0, 1)))])
    for interp in (pa.INTERP_REFERENCE, pa.INTERP_CONSISTENT):
        got = d.interpolate(pts, interp)
        for i in range(len(pts)):
            inside = orc.contains(o, pts[i])
            assert got[i, 3] == (1.0 if inside else 0.0), i
            if inside:
                want = orc.interpolate(o, pts[i], interp)
                assert got[i, :3].tobytes() == want.tobytes(), (interp, i, pts[i], got[i], want)


@pytest.mark.parametrize("h", [None, 2.0 ** -21], ids=["workload", "below"])
def test_rounds_fallbacks(ctx, orc, h):
    """Seeds on the node lattice nudged off their planes (remainders and
    lower-face coordinates below 2^-40 in the first steps), and a block whose
    spacing takes the division in every step: the reference round and the
    fused round equal the oracle's round bit for bit."""
    n = 25
    d, o, _ = single_block(ctx, orc, n, h=h)
    hs = o.spacing[0]
    dt = float(f32(0.3) * hs)  # about a third of a cell per step
    rng = np.random.default_rng(5)
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(2) * hs] * 3, 80, 0)
    pos = seeds["position"]
    nudge = rng.integers(0, len(TINY) + 1, size=(len(seeds), 3))
    for j, t in enumerate(TINY):
        pos[:, :3] = np.where(nudge == j, pos[:, :3] + t, pos[:, :3])
    seeds["position"] = pos
    _, oc, ocnt, _, osteps = orc.trace(seeds, o, dt=dt, integrator=orc.RK4)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(dt)
    parts = pa.Particles.from_host(ctx, seeds)
    ri = t.compute_round_info(parts)
    chunks = t.allocate(ri)
    res = t.trace(parts, chunks, ri)
    assert res["total_steps"] == osteps > 0
    assert np.array_equal(res["counts"].download(np.int32, len(seeds)), ocnt)
    assert chunks[0][0].download(np.float32, chunks[0][1] * 4).reshape(-1, 4).tobytes() == oc[0].tobytes()
    # the fused round on cell records (hybrid engine)
    d.set_layout(pa.LAYOUT_CELLS)
    t.set_engine(pa.ENGINE_HYBRID)
    want = orc.prune(oc[0])
    for _ in range(2):
        r = t.trace_union_round(parts, 80, 2, 0.75)
        assert r["total_steps"] == osteps and r["n_vertices"] == len(want)
        assert r["vertices"].download(np.float32, 4 * len(want)).tobytes() == want.tobytes()
        assert np.array_equal(r["indices"].download(np.int32, r["n_indices"]), orc.index_generate(want))
