"""Pins the CPU oracle (oracle/pa_oracle.cpp) against closed forms,
independent implementations (scipy, colorsys, numpy.linalg.eigh fixtures)
and brute force, so that the GPU parity tests compare against something
known to be right.  CPU only."""
import colorsys
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
f32 = np.float32


def uniform_field(orc, n, value, h=1.0, offset=(0.0, 0.0, 0.0), size=None):
    data = np.zeros((n, n, n, 3), np.float32)
    data[...] = np.asarray(value, np.float32)
    size = size if size is not None else [f32(n - 1) * f32(h)] * 3
    return orc.Field(data, offset, size, [h] * 3)


def particles(positions, R, vfi=-1):
    from oracle.oracle import PARTICLE_DTYPE

    p = np.zeros(len(positions), PARTICLE_DTYPE)
    p["position"][:, :3] = np.asarray(positions, np.float32)
    p["remaining_iterations"] = R
    p["vector_field_index"] = vfi
    return p


# ----------------------------------------------------------------- seeds
def test_seed_lattice_bruteforce(orc):
    for size, stride, off in [((4.0, 3.0, 2.0), (1.0, 1.0, 1.0), (0.5, -1.0, 2.0)),
                              ((1.0, 1.0, 1.0), (0.3, 0.25, 0.5), (0.0, 0.0, 0.0)),
                              ((2.0, 2.0, 2.0), (2.5, 1.0, 1.0), (0.0, 0.0, 0.0))]:
        n, dims = orc.seed_count(size, stride)
        expect_dims = tuple(int(f32(s) / f32(t)) for s, t in zip(size, stride))
        assert dims == expect_dims and n == int(np.prod(dims))
        p = orc.seed_generate(off, size, stride, 7, 3)
        assert len(p) == n
        for idx in range(n):
            i, j, k = np.unravel_index(idx, dims)  # C order, z fastest (index.hpp:9-19)
            want = [f32(off[a]) + f32(stride[a]) * f32(v) for a, v in enumerate((i, j, k))]
            assert list(p["position"][idx][:3]) == want
            assert p["position"][idx][3] == 0.0
            assert tuple(p["original_voxel"][idx]) == (i, j, k)
        assert (p["remaining_iterations"] == 7).all() and (p["vector_field_index"] == -1).all()
        assert (p["original_rank"] == 3).all()


def test_seed_config1_count(orc):
    h = f32(2 * np.pi / 63)
    n, dims = orc.seed_count([f32(64) * h] * 3, [f32(4) * h] * 3)
    assert dims == (16, 16, 16) and n == 4096


# ------------------------------------------------------- contains / interp
def test_interpolate_nodes_exact(orc):
    rng = np.random.default_rng(1)
    data = rng.normal(size=(6, 5, 7, 3)).astype(np.float32)
    F = orc.Field(data, (0, 0, 0), (2.5, 2.0, 3.0), (0.5, 0.5, 0.5))
    for i in range(5):
        for j in range(4):
            for k in range(6):
                p = np.array([0.5 * i, 0.5 * j, 0.5 * k, 0], np.float32)
                assert orc.contains(F, p)
                assert (orc.interpolate(F, p) == data[i, j, k]).all()


def test_interpolate_matches_scipy(orc):
    from scipy.interpolate import RegularGridInterpolator

    rng = np.random.default_rng(2)
    dims = (9, 7, 8)
    h = (0.37, 0.21, 0.5)
    off = (-1.0, 0.25, 3.0)
    data = rng.normal(size=dims + (3,))
    F64 = orc.Field64(data, off, [(d - 1) * s for d, s in zip(dims, h)], h)
    F32 = orc.Field(data.astype(np.float32), off, [f32((d - 1) * s) for d, s in zip(dims, h)], h)
    axes = [off[i] + h[i] * np.arange(dims[i]) for i in range(3)]
    interp = RegularGridInterpolator(axes, data, method="linear")
    for _ in range(500):
        p = np.array([off[i] + rng.uniform(0, (dims[i] - 1.0001) * h[i]) for i in range(3)] + [0.0])
        want = interp(p[:3])[0]
        got64 = np.zeros(3)
        import ctypes as C

        assert orc.lib().orc_contains_f64(F64.ref, p.ctypes.data_as(C.c_void_p)) == 1
        orc.lib().orc_interpolate_f64(F64.ref, p.ctypes.data_as(C.c_void_p), got64.ctypes.data_as(C.c_void_p))
        np.testing.assert_allclose(got64, want, rtol=0, atol=1e-12)
        got32 = orc.interpolate(F32, p.astype(np.float32))
        np.testing.assert_allclose(got32, want, rtol=0, atol=5e-5)


def test_interpolate_affine_field(orc):
    # Trilinear interpolation reproduces affine fields u = a + B x.
    rng = np.random.default_rng(3)
    n, h = 6, 0.25
    a = rng.normal(size=3)
    B = rng.normal(size=(3, 3))
    g = np.arange(n) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    data = a + X @ B.T
    F64 = orc.Field64(data, (0, 0, 0), [(n - 1) * h] * 3, [h] * 3)
    import ctypes as C

    for _ in range(200):
        p = np.array(list(rng.uniform(0, (n - 1) * h - 1e-9, size=3)) + [0.0])
        out = np.zeros(3)
        orc.lib().orc_interpolate_f64(F64.ref, p.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        np.testing.assert_allclose(out, a + B @ p[:3], atol=1e-12)


def test_contains_cell_rule(orc):
    F = uniform_field(orc, 5, (1, 0, 0))  # nodes 0..4, cells 0..3, h = 1
    assert orc.contains(F, [0, 0, 0, 0])
    assert orc.contains(F, [3.999, 0, 0, 0])
    assert not orc.contains(F, [4.0, 0, 0, 0])  # subscript 4 = shape-1 fails
    assert not orc.contains(F, [-1e-6, 0, 0, 0])  # negative subscript -> huge size_t
    assert orc.contains(F, [-0.0, 0, 0, 0])  # -0 -> subscript -0 -> 0
    assert not orc.contains(F, [np.nan, 0, 0, 0])
    assert not orc.contains(F, [0, 0, np.inf, 0])


# -------------------------------------------------------------------- step
def test_step_closed_forms(orc):
    rng = np.random.default_rng(4)
    for _ in range(200):
        p = np.array(list(rng.uniform(-10, 10, 3)) + [0.0], np.float32)
        v = rng.uniform(-3, 3, 3).astype(np.float32)
        dt = f32(rng.uniform(1e-3, 0.1))
        e = orc.step(orc.EULER, p, v, dt)
        assert (e[:3] == (p[:3] + dt * v)).all() and e[3] == 0.0  # one mul, one add
        r = orc.step(orc.RK4, p, v, dt)
        exact = p[:3].astype(np.float64) + float(dt) * v.astype(np.float64)
        ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
        assert (np.abs(r[:3] - exact) <= 4 * ulp + 1e-7).all()


# ------------------------------------------------------------- round info
def test_round_info_closed_forms(orc):
    p = particles(np.zeros((4096, 3)), 258)
    info = orc.round_info(p)
    ppic = 134217727 // 258
    assert info == (258, ppic * 258, 4096 * 258, 0, 1)
    p = particles(np.zeros((262144, 3)), 1002)
    ppic = 134217727 // 1002
    assert ppic == 133949
    info = orc.round_info(p, curve_offset=5)
    assert info == (1002, ppic * 1002, (262144 % ppic) * 1002, 5, 2)
    # N == ppic: the last chunk is empty
    p = particles(np.zeros((ppic, 3)), 1002)
    assert orc.round_info(p)[2] == 0 and orc.round_info(p)[4] == 2
    with pytest.raises(ValueError):
        orc.round_info(particles(np.zeros((0, 3)), 1))


# --------------------------------------------------------------- tracing
def test_trace_uniform_flow_euler_closed_form(orc):
    # u = (1,0,0), h = 1, dt = 3/8 exact: x_k = 0.25 + 0.375 k.  The first
    # uncontained vertex is x = 8.125 (k = 21) at iteration 22.
    F = uniform_field(orc, 9, (1, 0, 0), size=[8.0] * 3)
    p = particles([[0.25, 2.0, 3.0]], 100)
    info, chunks, counts, hand, steps = orc.trace(p, F, neighbor_rank=[-1, 7, -1, -1, -1, -1], dt=0.375,
                                                  integrator=orc.EULER)
    v = chunks[0]
    assert counts[0] == 23 and steps == 21
    np.testing.assert_array_equal(v[:22, 0], 0.25 + 0.375 * np.arange(22))
    assert (v[22] == orc.TERMINATION).all()
    assert (v[23:100] == orc.INVALID).all()
    assert len(hand) == 1 and hand[0]["kind"] == 0 and hand[0]["slot"] == 1
    hp = hand[0]["particle"]
    assert hp["remaining_iterations"] == 100 - 22 and hp["position"][0] == 8.125
    assert hp["vector_field_index"] == -1 and hp["original_rank"] == 0
    # Without a +x neighbour the particle is dropped.
    _, _, _, hand, _ = orc.trace(p, F, dt=0.375, integrator=orc.EULER)
    assert len(hand) == 0


def test_trace_cell_face_mismatch_drops(orc):
    # Exit exactly at x == max: the cell rule fails but x > max is false.
    F = uniform_field(orc, 9, (1, 0, 0), size=[8.0] * 3)
    p = particles([[0.5, 2.0, 3.0]], 100)
    _, _, counts, hand, steps = orc.trace(p, F, neighbor_rank=[1] * 6, dt=0.375, integrator=orc.EULER)
    assert steps == 20 and counts[0] == 22 and len(hand) == 0


def test_trace_lifetimes_and_stop(orc):
    F0 = uniform_field(orc, 6, (0, 0, 0))
    F1 = uniform_field(orc, 6, (1e-5, -1e-5, 0))  # isZero: every |v_i| <= 1e-5
    F2 = uniform_field(orc, 6, (2e-5, 0, 0))
    Rs = [0, 1, 2, 3, 10]
    p = particles([[1.0, 1.0, 1.0]] * len(Rs), 0)
    p["remaining_iterations"] = Rs
    for F, moves in ((F0, False), (F1, False), (F2, True)):
        _, _, counts, _, steps = orc.trace(p, F, dt=0.01)
        for i, R in enumerate(Rs):
            if R <= 1:
                assert counts[i] == 1  # seed only, no termination vertex
            elif moves:
                assert counts[i] == R  # seed + R-2 steps + termination
            else:
                assert counts[i] == 2  # stops at the first iteration
        assert steps == (sum(max(R - 2, 0) for R in Rs) if moves else 0)
    with pytest.raises(ValueError):
        orc.trace(particles([[1.0, 1.0, 1.0]], 0), F0)  # maximum R = 0: undefined in the reference


def test_trace_face_priority_and_neighbor_map(orc):
    F = uniform_field(orc, 5, (-1, -1, 0), size=[4.0] * 3)
    p = particles([[0.25, 0.25, 1.0]], 50)
    # exits through x<min and y<min at once: x first (slot 0) if present
    _, _, _, hand, _ = orc.trace(p, F, neighbor_rank=[10, 11, 12, 13, 14, 15], dt=0.5, integrator=orc.EULER)
    assert hand[0]["kind"] == 0 and hand[0]["slot"] == 0
    _, _, _, hand, _ = orc.trace(p, F, neighbor_rank=[-1, 11, 12, 13, 14, 15], dt=0.5, integrator=orc.EULER)
    assert hand[0]["slot"] == 2
    # a particle traced in neighbour block 3 goes to map 1, slot 3
    nb = [None, None, None, F, None, None]
    p = particles([[0.25, 0.25, 1.0]], 50, vfi=3)
    _, _, _, hand, _ = orc.trace(p, F, neighbors=nb, neighbor_rank=[10, 11, 12, 13, 14, 15], dt=0.5,
                                 integrator=orc.EULER)
    assert hand[0]["kind"] == 1 and hand[0]["slot"] == 3


def test_trace_nan_position(orc):
    F = uniform_field(orc, 5, (1, 0, 0))
    p = particles([[np.nan, 1.0, 1.0]], 10)
    _, chunks, counts, hand, steps = orc.trace(p, F, neighbor_rank=[1] * 6)
    assert counts[0] == 2 and steps == 0 and len(hand) == 0


def test_trace_linear_field_matches_matrix_power(orc):
    # Affine field u = A x sampled on a grid is reproduced exactly by
    # trilinear interpolation, so Euler gives x_n = (I + dt A)^n x_0.
    n, h, dt, steps = 33, 1.0 / 16, 0.01, 40
    A = np.array([[0.1, -0.3, 0.0], [0.3, 0.1, 0.05], [0.0, -0.05, -0.2]])
    g = np.arange(n) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    U = X @ A.T
    F64 = orc.Field64(U, (0, 0, 0), [(n - 1) * h] * 3, [h] * 3)
    F32 = orc.Field(U.astype(np.float32), (0, 0, 0), [f32((n - 1) * h)] * 3, [h] * 3)
    x0 = np.array([[1.0, 1.0, 1.0, 0.0], [0.9, 1.2, 1.1, 0.0]])
    e64, s64 = orc.endpoints_f64(x0, steps + 2, F64, dt=dt, integrator=orc.EULER)
    M = np.linalg.matrix_power(np.eye(3) + dt * A, steps)
    np.testing.assert_allclose(e64[:, :3], x0[:, :3] @ M.T, rtol=1e-12)
    assert (s64 == steps).all()
    e32, s32 = orc.endpoints(particles(x0[:, :3], steps + 2), F32, dt=dt, integrator=orc.EULER)
    np.testing.assert_allclose(e32[:, :3], x0[:, :3] @ M.T, rtol=2e-6)
    # The frozen-derivative RK4 is the same map up to rounding.
    e32r, _ = orc.endpoints(particles(x0[:, :3], steps + 2), F32, dt=dt, integrator=orc.RK4)
    np.testing.assert_allclose(e32r[:, :3], x0[:, :3] @ M.T, rtol=2e-6)
    # Textbook RK4 follows exp(T A) to fourth order.
    from scipy.linalg import expm

    e64c, _ = orc.endpoints_f64(x0, steps + 2, F64, dt=dt, integrator=orc.CLASSIC_RK4)
    np.testing.assert_allclose(e64c[:, :3], x0[:, :3] @ expm(steps * dt * A).T, rtol=1e-9)


def test_trace_slot_layout_and_fp64_twin(orc):
    from pars_amd import workloads as W

    n = 32
    h = 2 * np.pi / (n - 1)
    info = W.BlockInfo(0, (0, 0, 0), (0, 0, 0), (n, n, n))
    U = W.abc_block_f64(info, h)
    hf = f32(h)
    F32 = orc.Field(U.astype(np.float32), (0, 0, 0), [f32(n) * hf] * 3, [hf] * 3)
    F64 = orc.Field64(U, (0, 0, 0), [n * h] * 3, [h] * 3)
    seeds = orc.seed_generate((0, 0, 0), F32.size, [f32(3) * hf] * 3, 60, 0)
    ri, chunks, counts, hand, steps = orc.trace(seeds, F32, dt=0.01)
    v = chunks[0].reshape(len(seeds), 60, 4)
    for p in range(len(seeds)):
        c = counts[p]
        assert (v[p, 0] == seeds["position"][p]).all()
        if c >= 2:
            assert (v[p, c - 1] == orc.TERMINATION).all()
        assert (v[p, c:] == orc.INVALID).all()
        assert not (v[p, 1:c - 1] == orc.TERMINATION).all(axis=1).any()
    assert steps == int(np.maximum(counts - 2, 0).sum())
    # FP32 with consistent weights tracks the FP64 twin to rounding.
    e32, s32 = orc.endpoints(seeds, F32, dt=0.01, interp=orc.INTERP_CONSISTENT)
    e64, s64 = orc.endpoints_f64(seeds["position"].astype(np.float64), 60, F64, dt=0.01)
    same = s32 == s64
    assert same.mean() > 0.95
    err = np.abs(e32[same, :3] - e64[same, :3]).max()
    assert err < 1e-4
    # The reference weights deviate by a cell's worth where the defect hits.
    e32r, s32r = orc.endpoints(seeds, F32, dt=0.01)
    same = s32r == s64
    assert np.abs(e32r[same, :3] - e64[same, :3]).max() > 1e-3


def test_interpolate_reference_cell_weight_defect(orc):
    # z = 5*h in float: (z - 0)/h rounds to exactly 5.0, but the exact fmod is
    # h - ulp, so the reference reads cell 5 at weight ~1 = node 6; the
    # consistent interpolant returns node 5.
    n = 32
    hf = f32(2 * np.pi / (n - 1))
    rng = np.random.default_rng(7)
    data = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    F = orc.Field(data, (0, 0, 0), [f32(n) * hf] * 3, [hf] * 3)
    z = f32(5) * hf
    assert f32(z / hf) == 5.0 and f32(np.fmod(z, hf) / hf) > 0.9999
    p = np.array([0, 0, z, 0], np.float32)
    np.testing.assert_allclose(orc.interpolate(F, p), data[0, 0, 6], atol=1e-5)
    np.testing.assert_allclose(orc.interpolate(F, p, orc.INTERP_CONSISTENT), data[0, 0, 5], atol=1e-5)


def test_union_walk_three_blocks(orc):
    # Own block [0,8) with +x neighbour [8,16): uniform +x flow crosses once
    # and leaves the union at the far face of the neighbour.
    data = np.zeros((9, 9, 9, 3), np.float32)
    data[..., 0] = 1.0
    own = orc.Field(data, (0, 0, 0), (8, 8, 8), (1, 1, 1))
    plus = orc.Field(data, (8, 0, 0), (8, 8, 8), (1, 1, 1))
    nbrs = [None, plus, None, None, None, None]
    ranks = [-1, 5, -1, -1, -1, -1]
    masks = [0, 0b000011, 0, 0, 0, 0]  # +x block: neighbours across -x (own) and +x
    p = particles([[0.25, 2.0, 3.0]], 200)
    u = orc.trace_union(p, own, nbrs, ranks, masks, dt=0.375, integrator=orc.EULER)
    assert list(u["segment_block"]) == [-1, 1]
    L = u["segment_length"]
    assert L[0] == 23  # as in the single-block closed form
    seg2 = u["vertices"][u["segment_begin"][1]:u["segment_begin"][1] + L[1]]
    assert seg2[0, 0] == 8.125  # starts at the hand-off vertex
    # the second segment runs until x >= 16 : 8.125 + 0.375 k >= 16 -> k = 21
    assert L[1] == 23 and (seg2[-1] == orc.TERMINATION).all()
    assert u["total_steps"] == 21 + 21


# ------------------------------------------------------ prune/index/color
def test_prune_and_index_bruteforce(orc):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(500, 4)).astype(np.float32)
    v[:, 3] = 0
    mask = rng.uniform(size=500) < 0.3
    v[mask] = orc.INVALID
    v[rng.uniform(size=500) < 0.1] = orc.TERMINATION
    pr = orc.prune(v)
    np.testing.assert_array_equal(pr, v[~(v == orc.INVALID).all(1)])
    idx = orc.index_generate(pr)
    t = (pr == orc.TERMINATION).all(1)
    want = [k for k in range(len(pr) - 1) if not t[k] and not t[k + 1]]
    np.testing.assert_array_equal(idx, want)
    with pytest.raises(ValueError):
        orc.index_generate(np.zeros((0, 4), np.float32))


def test_index_uniform_lifetimes_closed_form(orc):
    # Every particle has R slots [seed, R-2 steps, T]: indices p*R + j, j < R-2.
    R, N = 7, 11
    v = np.ones((N * R, 4), np.float32)
    v[R - 1::R] = orc.TERMINATION
    idx = orc.index_generate(v)
    want = np.concatenate([p * R + np.arange(R - 2) for p in range(N)])
    np.testing.assert_array_equal(idx, want)


def test_color_axis_closed_forms(orc):
    seg = lambda d: np.array([[1, 1, 1, 0], [1 + d[0], 1 + d[1], 1 + d[2], 0]], np.float32)
    idx = np.array([0], np.int32)
    c = orc.color_generate(seg((0.5, 0, 0)), idx, 4)
    np.testing.assert_array_equal(c[0], [1, 0, 0, 1])
    c = orc.color_generate(seg((0, 0.5, 0)), idx, 4)
    np.testing.assert_array_equal(c[0], [0, 0, 1, 1])  # y and z swapped
    c = orc.color_generate(seg((0, 0, -2)), idx, 4)
    np.testing.assert_array_equal(c[0], [0, 1, 0, 1])
    for s in (0.0, 0.25, 0.75, 1.0):
        c = orc.color_generate(seg((0.5, 0, 0)), idx, 2, s)
        np.testing.assert_allclose(c[0], [1, 1 - s, 1 - s, 1], atol=1e-7)
    # zero-length segment: normalized() returns the zero vector
    c = orc.color_generate(seg((0, 0, 0)), idx, 4)
    np.testing.assert_array_equal(c[0], [0, 0, 0, 1])
    c = orc.color_generate(seg((0, 0, 0)), idx, 2, 0.5)
    assert np.isnan(c[0][:3]).all()


def _spherical_hue(t):
    s1 = math.atan2(t[1], t[0])
    if s1 < 0:
        s1 += math.pi
    if s1 >= math.pi:
        s1 -= math.pi
    s1 = (math.pi - s1) / math.pi
    s2 = math.acos(t[2] / math.sqrt(t[0] ** 2 + t[1] ** 2 + t[2] ** 2))
    if s2 >= math.pi / 2:
        s2 = math.pi - s2
    return s1, s2 / (math.pi / 2)


def test_color_hsv_hsl_against_colorsys(orc):
    rng = np.random.default_rng(6)
    for _ in range(300):
        d = rng.normal(size=3)
        v = np.array([[0, 0, 0, 0], [d[0], d[1], d[2], 0]], np.float32)
        t = (v[1, :3].astype(np.float64)) / np.linalg.norm(v[1, :3].astype(np.float64))
        h, s2 = _spherical_hue(t)
        h = min(h, 1.0 - 1e-12)  # colorsys wraps h == 1 to 0; the same colour
        param = float(rng.uniform(0, 1))
        got = orc.color_generate(v, np.array([0], np.int32), 2, param)[0]
        np.testing.assert_allclose(got[:3], colorsys.hsv_to_rgb(h, param, s2), atol=2e-6)
        got = orc.color_generate(v, np.array([0], np.int32), 3, param)[0]
        np.testing.assert_allclose(got[:3], colorsys.hsv_to_rgb(h, s2, param), atol=2e-6)
        got = orc.color_generate(v, np.array([0], np.int32), 0, param)[0]
        np.testing.assert_allclose(got[:3], colorsys.hls_to_rgb(h, s2, param), atol=2e-6)
        got = orc.color_generate(v, np.array([0], np.int32), 1, param)[0]
        np.testing.assert_allclose(got[:3], colorsys.hls_to_rgb(h, param, s2), atol=2e-6)


def test_color_curve_end_rule(orc):
    # two curves: indices 0,1 then 4; vertex 2 (end of curve 1) copies index 1's
    # colour, vertex 5 (end of the last index) stays unwritten
    v = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [1, 1, 0, 0], orc.TERMINATION, [0, 0, 0, 0], [0, 0, 1, 0]],
                 np.float32)
    idx = orc.index_generate(v)
    np.testing.assert_array_equal(idx, [0, 1, 4])
    c = orc.color_generate(v, idx, 4)
    np.testing.assert_array_equal(c[0], [1, 0, 0, 1])
    np.testing.assert_array_equal(c[1], [0, 0, 1, 1])
    np.testing.assert_array_equal(c[2], c[1])
    np.testing.assert_array_equal(c[3], [0, 0, 0, 0])
    np.testing.assert_array_equal(c[4], [0, 1, 0, 1])
    np.testing.assert_array_equal(c[5], [0, 0, 0, 0])


# ---------------------------------------------------------- advect / FTLE
def test_advect_closed_forms(orc):
    F = uniform_field(orc, 9, (1, 0, 0), size=[8.0] * 3)
    p = particles([[0.25, 2.0, 3.0], [0.25, 2.0, 3.0], [0.25, 2.0, 3.0]], 0)
    p["remaining_iterations"] = [100, 10, 1]
    out, state, slot, steps = orc.advect(p, F, dt=0.375, integrator=orc.EULER)
    # particle 0 leaves at iteration 22 (x = 8.125), no neighbour: inactive
    assert state[0] == 1 and out[0]["remaining_iterations"] == 0 and out[0]["position"][0] == 8.125
    # particle 1 does R-1 = 9 steps
    assert state[1] == 1 and out[1]["position"][0] == 0.25 + 9 * 0.375
    # R <= 1: never stepped, lost in the reference
    assert state[2] == 3 and out[2]["position"][0] == 0.25
    assert steps == 21 + 9
    out, state, slot, _ = orc.advect(p[:1], F, neighbor_rank=[-1, 4, -1, -1, -1, -1], dt=0.375,
                                     integrator=orc.EULER)
    assert state[0] == 2 and slot[0] == 1 and out[0]["remaining_iterations"] == 100 - 22


def test_eigh3_against_lapack_fixtures(orc):
    g = np.load(os.path.join(GOLDEN, "eigh3.npz"))
    for m, ev in zip(g["matrices"], g["eigenvalues"]):
        got = orc.eigh3(m)
        scale = np.abs(ev).max()
        np.testing.assert_allclose(got, ev, rtol=0, atol=1e-13 * scale + 1e-300)


def test_ftle_saddle_and_rotation(orc):
    n, h, lam, dt, steps = 12, 0.1, 0.7, 0.01, 50
    g = np.arange(n) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    T = steps * dt
    a, b = (1 + lam * dt) ** steps, (1 - lam * dt) ** steps
    fm = (X * np.array([a, b, 1.0])).astype(np.float32)
    ftle, lmax = orc.ftle(fm, [h] * 3, T)
    want = steps * math.log(1 + lam * dt) / T
    np.testing.assert_allclose(ftle, want, rtol=2e-6)
    # rotation: C = I, FTLE = 0
    th = 0.3
    Rm = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    fm = (X @ Rm.T).astype(np.float32)
    ftle, _ = orc.ftle(fm, [h] * 3, 1.0)
    assert np.abs(ftle).max() < 1e-5
    fa, _ = orc.ftle(fm, [h] * 3, 1.0, mode=1)
    assert np.abs(fa).max() < 1e-5


def test_flow_map_saddle_pipeline(orc):
    # u = (lam x, -lam y, 0) about the block centre; Euler flow map of the
    # interior gives FTLE = (R-1) ln(1 + lam dt) / T exactly in real arithmetic.
    n, h, lam, dt, R = 17, 1.0 / 16, 0.5, 0.01, 21
    g = np.arange(n) * h - 0.5
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    U = np.zeros_like(X)
    U[..., 0] = lam * X[..., 0]
    U[..., 1] = -lam * X[..., 1]
    F = orc.Field(U.astype(np.float32), (-0.5, -0.5, -0.5), [f32(16 * h)] * 3, [h] * 3)
    fm, steps = orc.flow_map(F, (16, 16, 16), R, 1.0, integrator=orc.EULER, dt=dt)
    T = (R - 1) * dt
    ftle, _ = orc.ftle(fm, [h] * 3, T)
    want = (R - 1) * math.log(1 + lam * dt) / T
    inner = ftle[4:12, 4:12, 4:12]
    np.testing.assert_allclose(inner, want, rtol=1e-4)
