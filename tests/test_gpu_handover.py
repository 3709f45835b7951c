"""Hand-overs from HBM: renderer chunks of a fused round, serialize records
of particles, and the surplus lists of load balancing, through the C ABI."""
import numpy as np
import pytest

import pars_amd as pa

from .helpers import ShareFields, f32

pytestmark = pytest.mark.gpu


def _round(ctx, orc, R=300):
    S = ShareFields(ctx, orc, 64, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    sh = S.share
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), R, 21)
    seeds["remaining_iterations"][::53] = 1   # curves without a termination vertex: pairs cross seeds
    seeds["remaining_iterations"][7::211] = 0
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    region = R + 40
    parts = pa.Particles.from_host(ctx, seeds)
    r = t.trace_union_round(parts, region, 2, 0.6, colors=ctx.upload(np.zeros(4 * len(seeds) * region, np.float32)))
    return seeds, region, r


@pytest.mark.parametrize("cap", [0, 20000, 3000])
def test_round_chunks(ctx, orc, cap):
    """Each chunk's index list is index_generator's on that chunk alone, its
    colours at the positions color_generator writes on that chunk are the
    stage's, no chunk exceeds the cap, and the chunks tile the round."""
    seeds, region, r = _round(ctx, orc)
    n, nv, ni = len(seeds), r["n_vertices"], r["n_indices"]
    verts = r["vertices"].download(np.float32, 4 * nv).reshape(nv, 4)
    cols_all = r["colors"].download(np.float32, 4 * nv).reshape(nv, 4)
    chunks = pa.round_chunks(ctx, r["counts"], n, region, r["vertices"], nv, r["indices"], ni, r["colors"],
                             max_vertices_per_chunk=cap)
    limit = cap if cap else 134217727
    ppic = limit // region
    assert len(chunks) == (n + ppic - 1) // ppic
    if cap:
        assert len(chunks) > 1
    counts = r["counts"].download(np.int32, n).astype(np.int64)
    v0 = 0
    for c, k in enumerate(chunks):
        assert k["first_particle"] == c * ppic and k["n_particles"] == min(ppic, n - c * ppic)
        assert k["n_vertices"] == counts[k["first_particle"]:k["first_particle"] + k["n_particles"]].sum()
        assert k["n_vertices"] <= limit
        assert k["vertices"] == r["vertices"].ptr + 16 * v0
        cv = verts[v0:v0 + k["n_vertices"]]
        want_i = orc.index_generate(cv)
        got_i = ctx_download_i32(ctx, k["indices"], k["n_indices"])
        assert np.array_equal(got_i, want_i), c
        want_c = orc.color_generate(cv, want_i, 2, 0.6)
        written = np.zeros(len(cv), bool)
        written[want_i] = True
        if len(want_i) > 1:
            ends = want_i[:-1][want_i[1:] != want_i[:-1] + 1]
            written[ends + 1] = True
        got_c = cols_all[v0:v0 + k["n_vertices"]]
        np.testing.assert_allclose(got_c[written], want_c[written], atol=1e-6, rtol=0)
        v0 += k["n_vertices"]
    assert v0 == nv


def ctx_download_i32(ctx, ptr, n):
    out = np.empty(max(n, 0), np.int32)
    if n:
        ctx.check(pa.lib().pa_gpu_memcpy_d2h(ctx.handle, out.ctypes.data_as(pa.C.c_void_p), pa.C.c_void_p(ptr),
                                             pa.C.c_size_t(4 * n)), "d2h")
    return out


def _random_particles(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, pa.PARTICLE_DTYPE)
    p["position"] = rng.normal(size=(n, 4)).astype(np.float32)
    p["position"][:, 3] = 0
    p["remaining_iterations"] = rng.integers(-5, 2000, n)
    p["vector_field_index"] = rng.integers(-1, 6, n)
    p["original_rank"] = rng.integers(0, 64, n)
    p["original_voxel"] = rng.integers(0, 512, (n, 3))
    return p


def test_particles_wire_roundtrip(ctx):
    """The records carry particle::serialize's fields in its order, 40
    bytes each (particle.hpp:16-28), and read back to the same particles."""
    p = _random_particles(5001, 1)
    p["position"][0] = [np.nan, -0.0, np.inf, 1e-40]
    dev = pa.Particles.from_host(ctx, p)
    wire = pa.particles_to_wire(ctx, dev)
    got = wire.download(pa.WIRE_DTYPE, len(p))
    want = np.zeros(len(p), pa.WIRE_DTYPE)
    for name in pa.WIRE_DTYPE.names:
        want[name] = p[name]
    assert got.tobytes() == want.tobytes()
    back = pa.particles_from_wire(ctx, wire, len(p)).download()
    assert back.tobytes() == p.tobytes()  # reserved words are zero in both


def test_load_balance_take(ctx):
    """Neighbour slots 0..5 in order take the last counts[k] particles of the
    list as it stands, none if that is more than it holds, with
    vector_field_index k ^ 1 (particle_tracer.cpp:168-184)."""
    p = _random_particles(1000, 2)
    dev = pa.Particles.from_host(ctx, p)
    for counts in ([3, 0, 5, 900, 2, 1], [0] * 6, [1000, 1, 0, 0, 0, 0], [400, 400, 300, 100, 100, 0]):
        wire, offs, keep = pa.load_balance_take(ctx, dev, counts)
        end, want, want_offs = len(p), [], [0]
        for k, c in enumerate(counts):
            if c > end:
                c = 0
            q = p[end - c:end].copy()
            q["vector_field_index"] = k ^ 1
            want.append(q)
            end -= c
            want_offs.append(want_offs[-1] + c)
        assert keep == end and offs == want_offs
        total = offs[6]
        got = wire.download(pa.WIRE_DTYPE, total) if total else np.zeros(0, pa.WIRE_DTYPE)
        w = np.zeros(total, pa.WIRE_DTYPE)
        allq = np.concatenate(want)
        for name in pa.WIRE_DTYPE.names:
            w[name] = allq[name]
        assert got.tobytes() == w.tobytes(), counts


def test_handoffs_to_wire(ctx, orc):
    """A reference round's hand-off lists, packed on the device, as serialize
    records."""
    S = ShareFields(ctx, orc, 64, 64, 21)
    sh = S.share
    t = S.tracer(ctx, 0.02, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_LANE)
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), 150, 21)
    parts = pa.Particles.from_host(ctx, seeds)
    info = t.compute_round_info(parts)
    chunks = t.allocate(info, fill_invalid=False)
    res = t.trace(parts, chunks, info)
    hand = res["handoff"]
    assert len(hand) > 0
    wire = pa.particles_to_wire(ctx, pa.Particles.from_host(ctx, hand))
    got = wire.download(pa.WIRE_DTYPE, len(hand))
    for name in pa.WIRE_DTYPE.names:
        assert np.array_equal(got[name], hand[name], equal_nan=True)
