"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs.  Integer, index and slot results are bit-exact; FP32 vertices
are compared bit-exactly as well (the GPU arithmetic is IEEE-identical to the
reference's), which implies the 1e-5 relative bar; colours from glibc vs
device atan2f/acosf are compared at 1e-6 absolute."""
import numpy as np
import pytest

import pars_amd as pa
from pars_amd import workloads as W

from .helpers import ShareFields, f32, oracle_integrator, single_block

pytestmark = pytest.mark.gpu

STORAGES = [pa.FP32, pa.BF16, pa.FP16]


def rel_ok(a, b, tol=1e-5):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= tol * np.maximum(np.abs(b), 1.0))


def test_seeds_bitexact(ctx, orc):
    cases = []
    h = f32(2 * np.pi / 63)
    cases.append(((0, 0, 0), [f32(64) * h] * 3, [f32(4) * h] * 3, 258, 0))
    h = f32(2 * np.pi / 511)
    o = f32(128) * h
    cases.append(((o, o, o), [f32(128) * h] * 3, [f32(2) * h] * 3, 1002, 21))
    cases.append(((0.5, -1.0, 2.0), (1.0, 2.0, 3.0), (0.3, 0.25, 0.5), 5, 3))
    for off, size, stride, R, rank in cases:
        d = pa.seed_generator.generate(ctx, off, size, stride, R, rank).download()
        o = orc.seed_generate(off, size, stride, R, rank)
        assert len(d) == len(o)
        assert d.tobytes() == o.tobytes()


@pytest.mark.parametrize("storage", STORAGES)
def test_field_storage_roundtrip(ctx, storage):
    rng = np.random.default_rng(11)
    data = (rng.normal(size=(9, 8, 7, 3)) * 3).astype(np.float32)
    data[0, 0, 0] = [1e-30, -0.0, 65504.0]
    f = pa.VectorField.from_host(ctx, data, (0, 0, 0), (8, 7, 6), (1, 1, 1), storage)
    got = f.read()
    want = W.quantize(data, storage)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert f.nbytes() == data.shape[0] * data.shape[1] * data.shape[2] * (16 if storage == pa.FP32 else 8)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("interp", [pa.INTERP_REFERENCE, pa.INTERP_CONSISTENT])
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
def test_interpolate_bitexact(ctx, orc, storage, interp, layout):
    d, o, _ = single_block(ctx, orc, 33, storage, layout=layout)
    rng = np.random.default_rng(12)
    h = o.spacing[0]
    n = 20000
    pts = np.zeros((n, 4), np.float32)
    pts[:, :3] = rng.uniform(-0.1, f32(33) * h + 0.1, size=(n, 3)).astype(np.float32)
    # node planes (where the reference's rounded quotient and exact remainder disagree)
    k = rng.integers(0, 33, size=(4000, 3))
    pts[:4000, :3] = k.astype(np.float32) * h
    pts[4000:4100, 0] = -0.0
    pts[4100:4110, 1] = np.nan
    got = d.interpolate(pts, interp)
    for i in range(n):
        inside = orc.contains(o, pts[i])
        assert got[i, 3] == (1.0 if inside else 0.0), i
        if inside:
            want = orc.interpolate(o, pts[i], interp)
            assert got[i, :3].tobytes() == want.tobytes(), (i, pts[i], got[i], want)


def _trace_both(ctx, orc, seeds_np, dev_tracer, orc_kwargs, local_orc):
    parts = pa.Particles.from_host(ctx, seeds_np)
    info = dev_tracer.compute_round_info(parts)
    oinfo = orc.round_info(seeds_np)
    assert (info.maximum_remaining_iterations, info.vertices_per_integral_curve,
            info.vertices_per_integral_curve_last, info.integral_curve_offset,
            info.integral_curve_count) == oinfo
    chunks = dev_tracer.allocate(info, fill_invalid=True)
    res = dev_tracer.trace(parts, chunks, info)
    o_info, o_chunks, o_counts, o_hand, o_steps = orc.trace(seeds_np, local_orc, **orc_kwargs)
    return chunks, res, o_chunks, o_counts, o_hand, o_steps


def test_trace_config1_bitexact(ctx, orc):
    """ABC 64^3, 4096 seeds (stride 4), R = 258, dt = 0.01, runge_kutta_4."""
    d, o, _ = single_block(ctx, orc, 64)
    h = o.spacing[0]
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(4) * h] * 3, 258, 0)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    chunks, res, oc, ocnt, ohand, osteps = _trace_both(ctx, orc, seeds, t, dict(dt=0.01, integrator=orc.RK4), o)
    got = chunks[0][0].download(np.float32, chunks[0][1] * 4).reshape(-1, 4)
    assert np.array_equal(res["counts"].download(np.int32, len(seeds)), ocnt)
    assert res["total_steps"] == osteps
    assert got.tobytes() == oc[0].tobytes()
    assert rel_ok(got, oc[0])
    assert res["handoff_counts"].sum() == 0 and len(ohand) == 0


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("integ", [pa.RUNGE_KUTTA_4, pa.EULER, pa.CLASSIC_RK4])
@pytest.mark.parametrize("interp", [pa.INTERP_REFERENCE, pa.INTERP_CONSISTENT])
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
@pytest.mark.parametrize("engine", [pa.ENGINE_COOPERATIVE, pa.ENGINE_LANE])
def test_trace_share_bitexact(ctx, orc, storage, integ, interp, layout, engine):
    """One rank's share (own block + 6 neighbours, 4x4x4 grid of a 64^3 ABC
    domain, rank 21), own-block seeds plus load-balanced particles that live in
    neighbour blocks; slots, counts and both hand-off maps compared."""
    S = ShareFields(ctx, orc, 64, 64, 21, storage, layout=layout)
    sh = S.share
    dt = 0.02
    t = S.tracer(ctx, dt, integ, interp, engine, waves=7 if engine == pa.ENGINE_COOPERATIVE else 0)
    own = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), 150, 21)
    extra = []
    for k, nb in enumerate(sh.neighbors):
        p = orc.seed_generate(nb.offset, nb.size, nb.spacing * f32(4), 120, nb.rank)
        p["vector_field_index"] = k
        extra.append(p)
    seeds = np.concatenate([own] + extra)
    oi = oracle_integrator(orc, integ)
    chunks, res, oc, ocnt, ohand, osteps = _trace_both(ctx, orc, seeds, t, S.oracle_kwargs(dt, oi, interp), S.orc[0])
    got = chunks[0][0].download(np.float32, chunks[0][1] * 4).reshape(-1, 4)
    assert np.array_equal(res["counts"].download(np.int32, len(seeds)), ocnt)
    assert res["total_steps"] == osteps and osteps > 0
    assert got.tobytes() == oc[0].tobytes()
    hc = res["handoff_counts"]
    want_hc = np.zeros((2, 6), np.int64)
    for h in ohand:
        want_hc[h["kind"], h["slot"]] += 1
    assert np.array_equal(hc, want_hc) and hc[0].sum() > 0 and hc[1].sum() > 0
    assert res["handoff"].tobytes() == ohand["particle"].tobytes()
    assert np.array_equal(res["handoff_sources"], ohand["source"])


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("region_extra", [64, 1])
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
@pytest.mark.parametrize("engine,order", [(pa.ENGINE_COOPERATIVE, "seed"), (pa.ENGINE_LANE, "seed"),
                                          (pa.ENGINE_LANE, "step"), (pa.ENGINE_QUAD, "chunked"),
                                          (pa.ENGINE_HYBRID, "group8")])
@pytest.mark.parametrize("integ", [pa.RUNGE_KUTTA_4, pa.CLASSIC_RK4])
def test_trace_union_bitexact(ctx, orc, storage, region_extra, layout, engine, order, integ):
    """Union-of-7 continuation vs the oracle's round emulation: every seed's
    segments, in order; a tiny region forces the continuation path.  Each
    pass is also packed by the counted prune of its slot order."""
    S = ShareFields(ctx, orc, 64, 64, 21, storage, layout=layout)
    sh = S.share
    dt, R = 0.02, 400
    t = S.tracer(ctx, dt, integ, pa.INTERP_REFERENCE, engine, waves=3 if engine == pa.ENGINE_COOPERATIVE else 0)
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), R, 21)
    ou = orc.trace_union(seeds, S.orc[0], S.orc[1:], sh.neighbor_ranks, sh.face_masks, dt=dt,
                         integrator=oracle_integrator(orc, integ))
    region = R + region_extra
    if order == "chunked":  # multiple of 8; the tiny case leaves no room for a second segment
        region = R if region_extra == 1 else R + 64
    n = len(seeds)
    pending = pa.Particles.from_host(ctx, seeds)
    seed_of = np.arange(n)
    got = {i: [] for i in range(n)}
    steps = 0
    for _ in range(64):
        r = t.trace_union(pending, region, order=order)
        steps += r["total_steps"]
        cnt = r["counts"].download(np.int32, pending.n)
        n8 = (pending.n + 7) // 8 * 8
        v = r["vertices"].download(np.float32, (n8 if order == "group8" else pending.n) * region * 4)
        if order == "group8":
            v = v.reshape(n8 // 8, region, 8, 4).transpose(0, 2, 1, 3).reshape(n8, region, 4)[:pending.n]
        elif order == "step":
            v = v.reshape(region, pending.n, 4).transpose(1, 0, 2)
        elif order == "chunked":  # [seed][chunk][component][8]
            v = v.reshape(pending.n, region // 8, 4, 8).transpose(0, 1, 3, 2).reshape(pending.n, region, 4)
        else:
            v = v.reshape(pending.n, region, 4)
        for j in range(pending.n):
            got[int(seed_of[j])].append(v[j, :cnt[j]])
        total = int(cnt.sum())
        if order == "step":
            packed, m, offs = pa.particle_tracer.prune_counted_step_major(ctx, r["vertices"].ptr, r["counts"],
                                                                          pending.n, total)
        elif order == "chunked":
            packed, m, offs = pa.particle_tracer.prune_counted_chunked(ctx, r["vertices"].ptr, region, r["counts"],
                                                                       pending.n, total)
        elif order == "group8":
            packed, m, offs = pa.particle_tracer.prune_counted_group8(ctx, r["vertices"].ptr, region, r["counts"],
                                                                      pending.n, total)
        else:
            packed, m, offs = pa.particle_tracer.prune_counted(ctx, r["vertices"].ptr, region, r["counts"],
                                                               pending.n, total)
        assert m == total
        want_packed = np.concatenate([v[j, :cnt[j]] for j in range(pending.n)])
        assert packed.download(np.float32, 4 * m).tobytes() == want_packed.tobytes()
        assert np.array_equal(offs.download(np.int64, pending.n + 1), np.concatenate([[0], np.cumsum(cnt)]))
        cont = r["continuation"]
        if cont.n == 0:
            break
        cp = cont.download()
        perm = np.argsort(cp["original_voxel"][:, 0], kind="stable")
        cp = cp[perm]
        seed_of = seed_of[cp["original_voxel"][:, 0]]
        pending = pa.Particles.from_host(ctx, cp)
    else:
        pytest.fail("continuation did not drain")
    assert steps == ou["total_steps"]
    for i in range(n):
        sel = np.nonzero(ou["segment_seed"] == i)[0]
        want = np.concatenate([ou["vertices"][ou["segment_begin"][s]:ou["segment_begin"][s] + ou["segment_length"][s]]
                               for s in sel])
        mine = np.concatenate(got[i])
        assert mine.tobytes() == want.tobytes(), i
    assert (ou["segment_block"] >= 0).any()  # some seeds did cross into neighbours


@pytest.mark.parametrize("n", [1, 13, 2049, 30001])
def test_union_round_fused_sizes(ctx, orc, n):
    """The fused finish runs as many workgroups as fit at once and hands the
    remaining 8-seed groups out through a queue: fewer groups than
    workgroups, a partial last group and many groups per workgroup all give
    the separate stages' vertices, indices and colours, twice in a row (the
    queue is reset by the round itself)."""
    S = ShareFields(ctx, orc, 128, 64, 21, pa.FP32, layout=pa.LAYOUT_CELLS)
    sh = S.share
    dt, R = 0.02, 120
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing, R, 21)  # one seed per node
    assert len(seeds) >= n
    seeds = seeds[:n]
    parts = pa.Particles.from_host(ctx, seeds)
    region = R + 40
    t = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    t2 = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)  # owns u's buffers
    u = t2.trace_union(parts, region, order="group8")
    total = int(u["counts"].download(np.int32, n).sum())
    packed, m, _ = pa.particle_tracer.prune_counted_group8(ctx, u["vertices"].ptr, region, u["counts"], n, total)
    idx, ni = pa.index_generator.generate(ctx, packed, m)
    # a vertex that the colour rule leaves alone keeps what the buffer held
    sentinel = np.full(n * region * 4, 7.5, np.float32)  # the round's vertex capacity
    col_sep = ctx.upload(sentinel)
    ctx.check(pa.lib().pa_gpu_color_generate(ctx.handle, pa.C.c_void_p(packed.ptr), pa.C.c_size_t(m),
                                             pa.C.c_void_p(idx.ptr), pa.C.c_size_t(ni), pa.C.c_int(2),
                                             pa.C.c_float(0.6), pa.C.c_void_p(col_sep.ptr)), "color")
    for _ in range(2):
        r = t.trace_union_round(parts, region, 2, 0.6, colors=ctx.upload(sentinel))
        assert r["continuation"].n == 0
        assert r["n_vertices"] == m and r["n_indices"] == ni
        assert r["vertices"].download(np.float32, 4 * m).tobytes() == packed.download(np.float32, 4 * m).tobytes()
        assert np.array_equal(r["indices"].download(np.int32, ni), idx.download(np.int32, ni))
        assert r["colors"].download(np.float32, 4 * m).tobytes() == col_sep.download(np.float32, 4 * m).tobytes()


@pytest.mark.parametrize("storage,layout", [(pa.FP32, pa.LAYOUT_CELLS), (pa.BF16, pa.LAYOUT_NODES),
                                            (pa.BF16, pa.LAYOUT_CELLS), (pa.FP16, pa.LAYOUT_CELLS)])
@pytest.mark.parametrize("mode", [2, 0, 4, 100])
def test_union_round_fused(ctx, orc, storage, layout, mode):
    """The fused round (union trace + prune + index + colour in one call)
    equals the separate stages on the same slots, bit for bit, including
    seeds that end without a termination vertex (remaining_iterations 1)."""
    S = ShareFields(ctx, orc, 64, 64, 21, storage, layout=layout)
    sh = S.share
    dt, R = 0.02, 300
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), R, 21)
    seeds["remaining_iterations"][::97] = 1  # curves of a single vertex: index pairs cross into the next seed
    seeds["remaining_iterations"][5::211] = 0
    n = len(seeds)
    parts = pa.Particles.from_host(ctx, seeds)
    region = R + 40
    t = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    sentinel = np.full(n * region * 4, 7.5, np.float32)
    col_fused = ctx.upload(sentinel)
    r = t.trace_union_round(parts, region, mode, 0.6, colors=col_fused)
    assert r["continuation"].n == 0
    # separate stages on the same slots
    t2 = S.tracer(ctx, dt, pa.RUNGE_KUTTA_4, pa.INTERP_REFERENCE, pa.ENGINE_HYBRID)
    u = t2.trace_union(parts, region, order="group8")
    assert u["total_steps"] == r["total_steps"]
    cnt = u["counts"].download(np.int32, n)
    assert np.array_equal(cnt, r["counts"].download(np.int32, n))
    total = int(cnt.sum())
    packed, m, _ = pa.particle_tracer.prune_counted_group8(ctx, u["vertices"].ptr, region, u["counts"], n, total)
    assert r["n_vertices"] == m == total
    assert r["vertices"].download(np.float32, 4 * m).tobytes() == packed.download(np.float32, 4 * m).tobytes()
    idx, ni = pa.index_generator.generate(ctx, packed, m)
    assert r["n_indices"] == ni
    assert np.array_equal(r["indices"].download(np.int32, ni), idx.download(np.int32, ni))
    col_sep = ctx.upload(sentinel)
    ctx.check(pa.lib().pa_gpu_color_generate(ctx.handle, pa.C.c_void_p(packed.ptr), pa.C.c_size_t(m),
                                             pa.C.c_void_p(idx.ptr), pa.C.c_size_t(ni), pa.C.c_int(mode),
                                             pa.C.c_float(0.6), pa.C.c_void_p(col_sep.ptr)), "color")
    assert col_fused.download(np.float32, 4 * m).tobytes() == col_sep.download(np.float32, 4 * m).tobytes()
    # and the index list is the oracle's on the packed vertices
    oidx = orc.index_generate(packed.download(np.float32, 4 * m).reshape(-1, 4))
    assert np.array_equal(idx.download(np.int32, ni), oidx)


def test_union_round_rejects_bad_particles(ctx, orc):
    """The fused round checks the particles on the device (the hybrid engine
    has no host pre-pass): a field index outside -1..5, a block that is not
    loaded, more iterations than region_slots and negative iterations each
    fail with their own status, and the next good call is unaffected."""
    d, o, _ = single_block(ctx, orc, 24)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.02)
    t.set_engine(pa.ENGINE_HYBRID)
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(4) * o.spacing[0]] * 3, 40, 0)
    region = 48
    bufs = t.union_buffers(len(seeds), region, endpoints=True)

    def run(p):
        return t.trace_union_round(pa.Particles.from_host(ctx, p), region, buffers=bufs)

    good = run(seeds)
    want_v = good["vertices"].download(np.float32, 4 * good["n_vertices"]).tobytes()
    cases = [("vector_field_index", 7, pa.E_INVALID), ("vector_field_index", 2, pa.E_STATE),
             ("remaining_iterations", region + 1, pa.E_CAPACITY), ("remaining_iterations", -3, pa.E_INVALID)]
    for key, value, status in cases:
        bad = seeds.copy()
        bad[key][len(bad) // 2] = value
        with pytest.raises(pa.PaError) as e:
            run(bad)
        assert e.value.status == status, (key, value)
    again = run(seeds)
    assert again["n_vertices"] == good["n_vertices"] and again["n_indices"] == good["n_indices"]
    assert again["vertices"].download(np.float32, 4 * again["n_vertices"]).tobytes() == want_v


def test_union_round_capacity(ctx, orc):
    """A fused round whose vertex or index capacity is too small fails with
    PA_E_CAPACITY, reports the needed sizes and writes nothing; the same
    call with room succeeds."""
    d, o, _ = single_block(ctx, orc, 24)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.02)
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(3) * o.spacing[0]] * 3, 40, 0)
    n, region = len(seeds), 48
    parts = pa.Particles.from_host(ctx, seeds)
    b = t.union_buffers(n, region, endpoints=True)
    good = t.trace_union_round(parts, region, buffers=b)
    nv, ni = good["n_vertices"], good["n_indices"]
    lib, C = pa.lib(), pa.C
    sentinel = np.full(4 * nv, 7.25, np.float32)
    for vcap, icap in ((nv - 1, nv), (nv, ni - 1)):
        verts, idx, cols = ctx.upload(sentinel), ctx.alloc(4 * nv), ctx.alloc(16 * nv)
        ncont, gv, gi, steps = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
        st = lib.pa_gpu_tracer_trace_union_round(
            t._h, C.c_void_p(parts.ptr), C.c_size_t(n), C.c_size_t(region), C.c_void_p(b["vertices"].ptr),
            C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.byref(ncont), C.c_int(2), C.c_float(0.75),
            C.c_void_p(verts.ptr), C.c_size_t(vcap), C.byref(gv), C.c_void_p(idx.ptr), C.c_size_t(icap), C.byref(gi),
            C.c_void_p(cols.ptr), C.c_void_p(b["endpoints"].ptr), C.c_void_p(b["seed_steps"].ptr), C.byref(steps))
        assert st == pa.E_CAPACITY
        assert (gv.value, gi.value) == (nv, ni)
        assert verts.download(np.float32, 4 * nv).tobytes() == sentinel.tobytes()


def test_hybrid_region_limit(ctx, orc):
    """The hybrid engine takes region_slots below 2^21: 2^21 fails with
    PA_E_INVALID before anything is written; at 2^21 - 1 the slots equal
    those of a small region."""
    d, o, _ = single_block(ctx, orc, 24, layout=pa.LAYOUT_CELLS)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.02)
    t.set_engine(pa.ENGINE_HYBRID)
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(12) * o.spacing[0]] * 3, 30, 0)
    n = len(seeds)
    assert 0 < n <= 8
    parts = pa.Particles.from_host(ctx, seeds)
    limit = (1 << 21) - 1
    small = t.trace_union(parts, 32, order="group8")
    want = small["vertices"].download(np.float32, 8 * 32 * 4).reshape(32, 8, 4)
    counts = small["counts"].download(np.int32, n)
    big = t.union_buffers(n, limit)
    sentinel = np.full(8 * 64 * 4, 5.5, np.float32)
    big["vertices"].upload(sentinel)
    with pytest.raises(pa.PaError) as e:
        t.trace_union(parts, limit + 1, buffers=big, order="group8")
    assert e.value.status == pa.E_INVALID
    assert big["vertices"].download(np.float32, sentinel.size).tobytes() == sentinel.tobytes()
    r = t.trace_union(parts, limit, buffers=big, order="group8")
    assert np.array_equal(r["counts"].download(np.int32, n), counts)
    got = r["vertices"].download(np.float32, 8 * 32 * 4).reshape(32, 8, 4)
    for p in range(n):
        assert got[:counts[p], p].tobytes() == want[:counts[p], p].tobytes()


def test_prune_index_color_parity(ctx, orc):
    d, o, _ = single_block(ctx, orc, 48)
    h = o.spacing[0]
    seeds = orc.seed_generate((0, 0, 0), o.size, [f32(3) * h] * 3, 120, 0)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.02)
    chunks, res, oc, ocnt, _, _ = _trace_both(ctx, orc, seeds, t, dict(dt=0.02, integrator=orc.RK4), o)
    buf, nslots = chunks[0]
    pruned, m = pa.particle_tracer.prune(ctx, buf, nslots)
    opr = orc.prune(oc[0])
    assert m == len(opr)
    assert pruned.download(np.float32, 4 * m).tobytes() == opr.tobytes()
    # counted prune of the same round equals the value prune
    packed, m2, offs = pa.particle_tracer.prune_counted(ctx, buf.ptr, 120, res["counts"], len(seeds))
    assert m2 == m and packed.download(np.float32, 4 * m).tobytes() == opr.tobytes()
    idx, ni = pa.index_generator.generate(ctx, pruned, m)
    oidx = orc.index_generate(opr)
    assert ni == len(oidx) and np.array_equal(idx.download(np.int32, ni), oidx)
    for mode in (0, 1, 2, 3, 4, 100):
        col = pa.color_generator.generate(ctx, pruned, m, idx, ni, mode, 0.6)
        got = col.download(np.float32, 4 * m).reshape(-1, 4)
        want = orc.color_generate(opr, oidx, mode, 0.6)
        if mode == 4:
            assert got.tobytes() == want.tobytes()
        else:
            np.testing.assert_allclose(got, want, atol=1e-6, rtol=0)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
def test_flow_map_and_ftle(ctx, orc, storage, layout):
    n, R, dt = 40, 101, 0.01
    d, o, _ = single_block(ctx, orc, n, storage, layout=layout)
    fmg = pa.flow_map_generator(ctx)
    fmg.set_vector_field(d)
    fmg.set_integrator("runge_kutta_4")
    fmg.set_step_size(dt)
    fm, dims, spacing, steps = fmg.generate((n, n, n), R, 1.0)
    ofm, osteps = orc.flow_map(o, (n, n, n), R, 1.0, integrator=orc.RK4, dt=dt)
    got = fm.download(np.float32, 3 * n ** 3).reshape(n, n, n, 3)
    assert steps == osteps
    assert got.tobytes() == ofm.tobytes()
    T = (R - 1) * dt
    for mode in (0, 1):
        ftg = pa.ftle_map_generator(ctx, mode)
        ftg.set_flow_map(fm, dims, spacing)
        ftg.set_time(T)
        out, lm = ftg.generate(want_lambda=True)
        g = out.download(np.float32, n ** 3).reshape(n, n, n)
        w, wl = orc.ftle(ofm, spacing, T, mode)
        assert np.all(np.abs(g.astype(np.float64) - w) <= 1e-4 * np.maximum(np.abs(w), 1e-3))
        glm = lm.download(np.float64, n ** 3).reshape(n, n, n)
        np.testing.assert_allclose(glm, wl, rtol=1e-9, atol=1e-12)


def test_advect_parity(ctx, orc):
    S = ShareFields(ctx, orc, 64, 64, 21)
    sh = S.share
    t = S.tracer(ctx, 0.02)
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), 300, 21)
    seeds["remaining_iterations"][::7] = 1
    parts = pa.Particles.from_host(ctx, seeds)
    state, slot, steps = t.advect(parts)
    out = parts.download()
    o_out, o_state, o_slot, o_steps = orc.advect(seeds, S.orc[0], neighbor_rank=sh.neighbor_ranks, dt=0.02,
                                                 integrator=orc.RK4)
    assert steps == o_steps
    assert np.array_equal(state, o_state) and np.array_equal(slot, o_slot)
    assert out.tobytes() == o_out.tobytes()
    assert (state == 2).any() and (state == 3).any()


def test_edge_cases(ctx, orc):
    d, o, _ = single_block(ctx, orc, 16)
    t = pa.particle_tracer(ctx)
    t.set_local_vector_field(d)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    empty = pa.Particles.from_host(ctx, np.zeros(0, pa.PARTICLE_DTYPE))
    with pytest.raises(pa.PaError) as e:
        t.compute_round_info(empty)
    assert e.value.status == pa.E_INVALID
    with pytest.raises(pa.PaError) as e:
        t.set_integrator("runge_kutta_dormand_prince_5")
    assert e.value.status == pa.E_UNSUPPORTED
    # R in {0,1,2,3}, NaN and outside positions
    p = np.zeros(6, pa.PARTICLE_DTYPE)
    p["position"][:, :3] = 1.0
    p["remaining_iterations"] = [0, 1, 2, 3, 9, 9]
    p["position"][4, 0] = np.nan
    p["position"][5, 0] = -5.0
    p["vector_field_index"] = -1
    chunks, res, oc, ocnt, ohand, osteps = _trace_both(ctx, orc, p, t, dict(dt=0.01, integrator=orc.RK4), o)
    assert np.array_equal(res["counts"].download(np.int32, 6), ocnt)
    assert list(ocnt) == [1, 1, 2, 3, 2, 2]
    got = chunks[0][0].download(np.float32, chunks[0][1] * 4).reshape(-1, 4)
    assert np.array_equal(got, oc[0], equal_nan=True)
    # a particle naming a neighbour block that is not loaded is rejected
    q = p.copy()
    q["vector_field_index"][0] = 2
    with pytest.raises(pa.PaError) as e:
        t.compute_round_info(pa.Particles.from_host(ctx, q))
    assert e.value.status == pa.E_STATE
    # index generation of an empty array is rejected (size()-1 underflow)
    with pytest.raises(pa.PaError):
        pa.index_generator.generate(ctx, ctx.alloc(16), 0)


def _pool_fields(ctx, orc, storage, layout):
    """Two fields of different size sharing storage and layout: ABC on 40^3 and
    a random-Fourier curl field on 48^3 (synthesized on the device)."""
    from pars_amd import workloads as W
    d0, o0, _ = single_block(ctx, orc, 40, storage, layout=layout)
    n = 48
    h = 2 * np.pi / (n - 1)
    hf = f32(h)
    modes = W.fourier_modes(12, 7)
    off = np.zeros(3, np.float32)
    size = np.array([f32(n) * hf] * 3, np.float32)
    sp = np.array([hf] * 3, np.float32)
    d1 = pa.VectorField.synthesize(ctx, "fourier", modes, (n, n, n), (0, 0, 0), h, off, size, sp,
                                   storage).set_layout(layout)
    o1 = orc.Field(d1.read(), off, size, sp)
    return [d0, d1], [o0, o1]


def _pool_particles(orc, fields_o, n_per, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for k, o in enumerate(fields_o):
        p = np.zeros(n_per, orc.PARTICLE_DTYPE)
        p["position"][:, :3] = rng.uniform(0.05, 0.95, (n_per, 3)).astype(np.float32) * o.size[0]
        p["remaining_iterations"] = rng.integers(0, 260, n_per)
        p["remaining_iterations"][:3] = [0, 1, 2]
        p["vector_field_index"] = k
        p["original_rank"] = k  # request id = field
        parts.append(p)
    allp = np.concatenate(parts)
    return allp[rng.permutation(len(allp))]


@pytest.mark.parametrize("storage,layout", [(pa.FP32, pa.LAYOUT_NODES), (pa.BF16, pa.LAYOUT_CELLS),
                                            (pa.FP32, pa.LAYOUT_CELLS), (pa.FP16, pa.LAYOUT_CELLS)])
def test_pool_batched_parity(ctx, orc, storage, layout):
    """Batched seed sets over two fields in one persistent kernel: every
    endpoint and step count equals the oracle's one-block trace; bounded
    passes and a pool holding one set alone give the same bits."""
    fields, fo = _pool_fields(ctx, orc, storage, layout)
    parts = _pool_particles(orc, fo, 700, 3)
    n = len(parts)
    t = pa.particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    pool = pa.ParticlePool(ctx, n, requests=2)
    dev = pa.Particles.from_host(ctx, parts)
    pool.init(dev, 2)
    total = pool.run(t, fields)
    res = pool.results()
    assert (res["status"] == pa.POOL_DONE).all()
    want_pos = np.zeros((n, 4), np.float32)
    want_steps = np.zeros(n, np.int32)
    for k in range(2):
        sel = np.nonzero(parts["vector_field_index"] == k)[0]
        q = parts[sel].copy()
        q["vector_field_index"] = -1
        e, st = orc.endpoints(q, fo[k], dt=0.01, integrator=orc.RK4)
        want_pos[sel], want_steps[sel] = e, st
    assert res["position"].tobytes() == want_pos.tobytes()
    assert np.array_equal(res["steps"], want_steps)
    assert total == int(want_steps.sum())
    done = pool.request_done.download(np.int32, 2)
    assert np.array_equal(done, np.bincount(parts["vector_field_index"], minlength=2))
    # bounded passes of 7 iterations reach the same state
    pool2 = pa.ParticlePool(ctx, n)
    pool2.init(dev, 2)
    passes, total2 = 0, 0
    while True:
        total2 += pool2.run(t, fields, max_iterations=7, waves_per_cu=2)
        passes += 1
        r2 = pool2.results()
        if (r2["status"] == pa.POOL_DONE).all():
            break
        assert passes < 200
    assert passes > 5 and total2 == total
    assert r2["position"].tobytes() == res["position"].tobytes()
    # one set alone
    sel = np.nonzero(parts["vector_field_index"] == 1)[0]
    pool3 = pa.ParticlePool(ctx, len(sel))
    pool3.init(pa.Particles.from_host(ctx, parts[sel]), 2)
    pool3.run(t, fields)
    assert pool3.results()["position"].tobytes() == res["position"][sel].tobytes()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200])
@pytest.mark.parametrize("waves", [0, 1, 64])
def test_pool_small_and_oversubscribed(ctx, orc, n, waves):
    """Pools smaller than, equal to and just past one wave's 64-entry claim,
    on the default launch, one wave per CU and more waves than stay
    resident (late workgroups find the counter past the end): every entry
    ends with the oracle's endpoint and step count, and each request's
    counter counts its entries once."""
    fields, fo = _pool_fields(ctx, orc, pa.BF16, pa.LAYOUT_CELLS)
    parts = _pool_particles(orc, fo, 100, 17)[:n]
    t = pa.particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    pool = pa.ParticlePool(ctx, n, requests=2)
    pool.init(pa.Particles.from_host(ctx, parts), 2)
    total = pool.run(t, fields, waves_per_cu=waves)
    res = pool.results()
    assert (res["status"] == pa.POOL_DONE).all()
    want_pos = np.zeros((n, 4), np.float32)
    want_steps = np.zeros(n, np.int32)
    for k in range(2):
        sel = np.nonzero(parts["vector_field_index"] == k)[0]
        if len(sel) == 0:
            continue
        q = parts[sel].copy()
        q["vector_field_index"] = -1
        want_pos[sel], want_steps[sel] = orc.endpoints(q, fo[k], dt=0.01, integrator=orc.RK4)
    assert res["position"].tobytes() == want_pos.tobytes()
    assert np.array_equal(res["steps"], want_steps) and total == int(want_steps.sum())
    assert np.array_equal(pool.request_done.download(np.int32, 2),
                          np.bincount(parts["original_rank"], minlength=2))


def test_pool_rejects_bad_field_index(ctx, orc):
    fields, fo = _pool_fields(ctx, orc, pa.FP32, pa.LAYOUT_NODES)
    parts = _pool_particles(orc, fo, 10, 4)
    parts["vector_field_index"][3] = 2
    pool = pa.ParticlePool(ctx, len(parts))
    with pytest.raises(pa.PaError):
        pool.init(pa.Particles.from_host(ctx, parts), 2)


def test_pool_request_ids_checked(ctx, orc):
    """Request ids index request_done: without request counters the pool
    reads no id (original_rank 3 is ignored and nothing is counted); with 2
    counters an id of 3 or -1 is rejected and marks only its entry invalid;
    a run given fewer fields than an entry names rejects that entry without
    stepping it and still finishes the others."""
    fields, fo = _pool_fields(ctx, orc, pa.FP32, pa.LAYOUT_NODES)
    t = pa.particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    parts = _pool_particles(orc, fo, 40, 9)
    parts["original_rank"] = 3
    n = len(parts)
    dev = pa.Particles.from_host(ctx, parts)
    plain = pa.ParticlePool(ctx, n)
    plain.init(dev, 2)
    total = plain.run(t, fields)
    assert (plain.results()["status"] == pa.POOL_DONE).all() and total > 0
    counted = pa.ParticlePool(ctx, n, requests=2)
    for bad in (3, -1):
        q = parts.copy()
        q["original_rank"] = 1
        q["original_rank"][[5, 17]] = bad
        with pytest.raises(pa.PaError) as e:
            counted.init(pa.Particles.from_host(ctx, q), 2)
        assert e.value.status == pa.E_INVALID
        st = counted.status.download(np.int32, n)
        assert st[5] == st[17] == pa.POOL_INVALID and (np.delete(st, [5, 17]) == pa.POOL_ACTIVE).all()
        counted.request_done.upload(np.zeros(2, np.int32))
        counted.run(t, fields)
        done = counted.request_done.download(np.int32, 2)
        assert done[0] == 0 and done[1] == n - 2
    # a run that is given only field 0: entries on field 1 are rejected, the rest finish
    one = pa.ParticlePool(ctx, n)
    one.init(dev, 2)
    with pytest.raises(pa.PaError) as e:
        one.run(t, fields[:1])
    assert e.value.status == pa.E_INVALID
    r = one.results()
    on1 = parts["vector_field_index"] == 1
    assert (r["status"][on1] == pa.POOL_INVALID).all() and (r["status"][~on1] == pa.POOL_DONE).all()
    assert (r["steps"][on1] == 0).all()
    assert r["position"][~on1].tobytes() == plain.results()["position"][~on1].tobytes()


@pytest.mark.parametrize("n_requests", [1, 7, 4096, 4097, 6000])
@pytest.mark.parametrize("layout", [pa.LAYOUT_NODES, pa.LAYOUT_CELLS])
def test_pool_request_counts(ctx, orc, n_requests, layout):
    """request_done holds, per request id, the number of its entries that
    finished: counted per workgroup in LDS and added once (up to 4,096
    requests) or added per particle (more), exact either way, over bounded
    passes and resident-wave counts that spread a request over many
    workgroups."""
    fields, fo = _pool_fields(ctx, orc, pa.BF16, layout)
    t = pa.particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(0.01)
    parts = _pool_particles(orc, fo, 3000, 11)
    rng = np.random.default_rng(n_requests)
    parts["original_rank"] = rng.integers(0, n_requests, len(parts))
    if n_requests > 3:  # a hot request: a third of the entries
        parts["original_rank"][rng.random(len(parts)) < 0.33] = n_requests - 1
    n = len(parts)
    want = np.bincount(parts["original_rank"], minlength=n_requests)
    dev = pa.Particles.from_host(ctx, parts)
    pool = pa.ParticlePool(ctx, n, requests=n_requests)
    for waves, budget in ((0, 0), (4, 37)):
        pool.request_done.upload(np.zeros(n_requests, np.int32))
        pool.init(dev, 2)
        passes = 0
        while True:
            pool.run(t, fields, max_iterations=budget, waves_per_cu=waves)
            passes += 1
            done = pool.request_done.download(np.int32, n_requests)
            if (pool.status.download(np.int32, n) == pa.POOL_DONE).all():
                break
            assert budget and passes < 20 and (done <= want).all()
        assert np.array_equal(done, want), (waves, budget)


def test_serving_alone_equals_mix(ctx, orc):
    """Continuous batching: requests that arrive while others run share the
    pool passes, and every request's endpoints equal its result alone."""
    from pars_amd import serve as S
    grid = 64
    field = S.make_field(ctx, pa.BF16, grid=grid, n_modes=8)
    reqs = S.make_requests(12, rate=20000.0, grid=grid, steps=120, size_range=(64, 4000))
    server = S.Server(ctx, [field], capacity=30000, pass_iterations=16, max_in_flight=64)
    steps, wall = server.serve(reqs)
    assert all(np.isfinite(r.done) for r in reqs)
    assert steps == sum(r.steps for r in reqs)
    overlapped = sum(1 for a in reqs for b in reqs if a is not b and a.admitted < b.admitted < a.done)
    assert overlapped > 0  # some requests joined while others were on the GPU
    alone = [S.Request(r.rid, 0.0, r.particles) for r in reqs]
    S.serve_one_at_a_time(ctx, [field], alone)
    for r, a in zip(reqs, alone):
        assert r.endpoints.tobytes() == a.endpoints.tobytes(), r.rid
        assert r.steps == a.steps
    # and the pool's one-block trace is the oracle's on the decoded field
    o = orc.Field(field.read(), field.offset, field.size, field.spacing)
    q = reqs[0].particles.copy()
    q["vector_field_index"] = -1
    e, st = orc.endpoints(q, o, dt=S.DT, integrator=orc.RK4)
    assert reqs[0].endpoints.tobytes() == e.tobytes()


def test_device_eigen_solver_on_lapack_fixtures(ctx):
    """The FTLE's FP64 symmetric 3x3 eigen solve (Smith's closed form) on the
    committed LAPACK (numpy.linalg.eigh) fixtures: near-degenerate,
    rank-deficient and 1e-6..1e6-scaled matrices included.  Error relative to
    the largest |eigenvalue| of each matrix."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eigh3.npz"))
    A, want = d["matrices"], d["eigenvalues"]
    sym = np.ascontiguousarray(np.stack([A[:, 0, 0], A[:, 0, 1], A[:, 0, 2], A[:, 1, 1], A[:, 1, 2], A[:, 2, 2]], 1))
    dev = ctx.upload(sym.astype(np.float64))
    out = ctx.alloc(len(A) * 24)
    ctx.check(pa.lib().pa_gpu_sym3_eigenvalues(ctx.handle, pa.C.c_void_p(dev.ptr), pa.C.c_size_t(len(A)),
                                               pa.C.c_void_p(out.ptr)), "eig")
    got = out.download(np.float64, 3 * len(A)).reshape(-1, 3)
    scale = np.maximum(np.abs(want).max(1), 1e-300)
    err = (np.abs(got - want).max(1) / scale)
    print("eigen solver: max %.3g, median %.3g relative to the largest |eigenvalue|" % (err.max(), np.median(err)))
    assert err.max() <= 1e-7
    assert np.median(err) <= 1e-13
    assert np.all(np.diff(got, axis=1) >= -1e-12 * scale[:, None])  # ascending


@pytest.mark.parametrize("storage,layout", [(pa.BF16, pa.LAYOUT_NODES), (pa.BF16, pa.LAYOUT_CELLS),
                                            (pa.FP32, pa.LAYOUT_CELLS)])
@pytest.mark.parametrize("capacity", [60000, 4500])
def test_resident_serving_alone_equals_mix(ctx, orc, capacity, storage, layout):
    """The device-resident serving loop: requests published while others run
    overlap on the device, a small ring admits them as space frees up, and
    every request's endpoints and steps equal its result alone and the
    oracle's (node bricks, and cell records kept per lane between steps)."""
    from pars_amd import serve as S
    grid = 64
    field = S.make_field(ctx, storage, grid=grid, n_modes=8).set_layout(layout)
    reqs = S.make_requests(14, rate=20000.0, grid=grid, steps=120, size_range=(64, 4000))
    server = S.ResidentServer(ctx, [field], capacity=capacity, request_slots=8, waves_per_cu=4)
    steps, wall = server.serve(reqs)
    total, launches = server.server.stop()
    server.close()
    assert all(np.isfinite(r.done) for r in reqs)
    assert steps == sum(r.steps for r in reqs) == total and launches >= 1
    overlapped = sum(1 for a in reqs for b in reqs if a is not b and a.admitted < b.admitted < a.done)
    assert overlapped > 0
    alone = [S.Request(r.rid, 0.0, r.particles) for r in reqs]
    S.serve_one_at_a_time(ctx, [field], alone)
    for r, a in zip(reqs, alone):
        assert r.endpoints.tobytes() == a.endpoints.tobytes(), r.rid
        assert r.steps == a.steps
    o = orc.Field(field.read(), field.offset, field.size, field.spacing)
    q = reqs[3].particles.copy()
    q["vector_field_index"] = -1
    e, st = orc.endpoints(q, o, dt=S.DT, integrator=orc.RK4)
    assert reqs[3].endpoints.tobytes() == e.tobytes()
