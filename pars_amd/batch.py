"""Batched uneven work on the particle pool, and its autotuner.

Many seed sets over fields of uneven size run in one persistent work-queue
kernel (pa_gpu_pool_run).  The workload generator builds the synthetic
batch of random-Fourier curl fields and seed sets; the autotuner times
candidate launch shapes (field layout: node bricks or cell records; resident
waves per CU) on a sample of the batch and keeps the fastest.
"""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import (BF16, LAYOUT_CELLS, LAYOUT_NODES, Context, ParticlePool, Particles, VectorField, particle_tracer,
               PARTICLE_DTYPE)
from . import workloads as W

GRIDS = (64, 96, 128, 160, 192, 224, 256, 256)
N_MODES = 48
N_SETS = 64
SET_SIZE = (1_000, 1_000_000)     # log-uniform
STEPS = (100, 2000)               # integrator steps per set, uniform
DT = 0.01


@dataclass
class BatchWorkload:
    """Host description of the batch: fields (by grid) and seed sets."""
    grids: Sequence[int]
    modes: List[np.ndarray]
    sets: List[dict]                       # field, size, steps
    particles: np.ndarray                  # all seeds, set by set; vfi = field, original_rank = set
    h: List[float] = field(default_factory=list)

    @property
    def n(self) -> int:
        return len(self.particles)

    @property
    def max_steps(self) -> int:
        return int(sum(s["size"] * s["steps"] for s in self.sets))


def make_workload(seed: int = W.DEFAULT_SEED, grids=GRIDS, n_sets=N_SETS, n_modes=N_MODES, size_range=SET_SIZE,
                  steps_range=STEPS, locality: bool = True) -> BatchWorkload:
    """Fields u = curl A with n_modes random modes each (fourier_modes), on
    grids^3 nodes over [0, 2 pi]^3; n_sets seed sets, sizes log-uniform in
    size_range, each on a uniformly chosen field with positions uniform in
    its domain, steps uniform in steps_range (R = steps + 2).  With
    `locality`, the seeds are stored in locality_order (results are per
    particle, so the order changes only the speed)."""
    rng = W.SplitMix64(seed)
    modes = [W.fourier_modes(n_modes, seed + 1 + k) for k in range(len(grids))]
    hs = [2.0 * math.pi / (g - 1) for g in grids]
    sets, chunks = [], []
    lo, hi = math.log(size_range[0]), math.log(size_range[1])
    for s in range(n_sets):
        size = int(round(math.exp(lo + (hi - lo) * rng.uniform())))
        fid = rng.randint(0, len(grids) - 1)
        steps = rng.randint(steps_range[0], steps_range[1])
        sets.append(dict(field=fid, size=size, steps=steps))
        g = grids[fid]
        hf = np.float32(hs[fid])
        extent = np.float32(g - 1) * hf  # the last node; positions stay inside the cells
        p = np.zeros(size, PARTICLE_DTYPE)
        # positions: a numpy generator seeded from the SplitMix stream
        u = np.random.default_rng(rng.next_u64() & 0xFFFFFFFF).uniform(0.0, 1.0, (size, 3))
        p["position"][:, :3] = (u * float(extent)).astype(np.float32)
        p["remaining_iterations"] = steps + 2
        p["vector_field_index"] = fid
        p["original_rank"] = s
        chunks.append(p)
    particles = np.concatenate(chunks)
    if locality:
        particles = particles[locality_order(particles, hs)]
    return BatchWorkload(grids=grids, modes=modes, sets=sets, particles=particles, h=hs)


def _spread10(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64) & np.uint64(0x3FF)
    out = np.zeros_like(x)
    for b in range(10):
        out |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def locality_order(particles: np.ndarray, hs: Sequence[float]) -> np.ndarray:
    """Permutation sorting particles by field, then by the Morton code of
    their start cell: neighbouring lanes of the pool kernel then gather
    neighbouring cells of the same field (shared cache lines)."""
    fid = particles["vector_field_index"].astype(np.int64)
    h = np.asarray(hs, np.float64)[fid]
    cell = np.clip(np.floor(particles["position"][:, :3] / h[:, None]), 0, 1023).astype(np.uint64)
    code = (_spread10(cell[:, 0]) << np.uint64(2)) | (_spread10(cell[:, 1]) << np.uint64(1)) | _spread10(cell[:, 2])
    key = (fid.astype(np.uint64) << np.uint64(30)) | code
    return np.argsort(key, kind="stable")


def build_fields(ctx: Context, wl: BatchWorkload, storage=BF16, layout=LAYOUT_NODES) -> List[VectorField]:
    out = []
    for g, m, h in zip(wl.grids, wl.modes, wl.h):
        hf = np.float32(h)
        size = np.array([np.float32(g) * hf] * 3, np.float32)
        f = VectorField.synthesize(ctx, "fourier", m, (g, g, g), (0, 0, 0), h, np.zeros(3, np.float32), size,
                                   np.array([hf] * 3, np.float32), storage)
        out.append(f.set_layout(layout))
    return out


def make_tracer(ctx: Context, dt: float = DT) -> particle_tracer:
    t = particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(dt)
    return t


@dataclass
class Tuning:
    layout: int
    waves_per_cu: int
    sample_steps_per_s: float
    table: Dict[str, float]


def autotune(ctx: Context, wl: BatchWorkload, storage=BF16, layouts=(LAYOUT_NODES, LAYOUT_CELLS),
             waves=(4, 8, 16, 0), sample_stride: int = 16, cache: Optional[str] = None) -> Tuning:
    """Time every (layout, waves per CU; 0 = as many as stay resident) on
    every sample_stride-th seed of the batch, best of two runs, and return
    the fastest; a full-occupancy launch within 3% of it is preferred (the
    sample's short runs understate what occupancy gives the whole batch:
    16 waves per CU won on the sample at 54.3 against 53.8 G steps/s and ran
    the batch in 33.8 ms against 31.6).  With `cache`, a previous result for
    the same device and batch shape is reused."""
    key = "%s|%d|%d|%d" % (ctx.device_info()["name"].strip(), wl.n, len(wl.sets), storage)
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        if key in known:
            k = known[key]
            return Tuning(k["layout"], k["waves_per_cu"], k["sample_steps_per_s"], k["table"])
    sample = wl.particles[::sample_stride].copy()
    dev = Particles.from_host(ctx, sample)
    t = make_tracer(ctx)
    table = {}
    best = None
    for lay in layouts:
        fields = build_fields(ctx, wl, storage, lay)
        pool = ParticlePool(ctx, len(sample))
        for wv in waves:
            pool.init(dev, len(fields))
            pool.run(t, fields, waves_per_cu=wv)  # warm
            rate = 0.0
            for _ in range(2):
                pool.init(dev, len(fields))
                ctx.synchronize()
                t0 = time.perf_counter()
                steps = pool.run(t, fields, waves_per_cu=wv)
                rate = max(rate, steps / (time.perf_counter() - t0))
            table["%s/%s" % ("cells" if lay == LAYOUT_CELLS else "nodes", "wmax" if wv == 0 else "w%d" % wv)] = rate
            if best is None or rate > best[2]:
                best = (lay, wv, rate)
        for f in fields:
            f.release()
    full = table.get("%s/wmax" % ("cells" if best[0] == LAYOUT_CELLS else "nodes"))
    if full is not None and full >= 0.97 * best[2]:
        best = (best[0], 0, full)
    tun = Tuning(best[0], best[1], best[2], table)
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[key] = dict(layout=tun.layout, waves_per_cu=tun.waves_per_cu,
                          sample_steps_per_s=tun.sample_steps_per_s, table=tun.table)
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return tun


def run_batch(ctx: Context, wl: BatchWorkload, fields: List[VectorField], waves_per_cu: int, pool=None, dev=None):
    """The whole batch in one persistent launch; returns (steps, seconds, pool)."""
    t = make_tracer(ctx)
    pool = ParticlePool(ctx, wl.n, requests=len(wl.sets)) if pool is None else pool
    dev = Particles.from_host(ctx, wl.particles) if dev is None else dev
    pool.request_done.upload(np.zeros(len(wl.sets), np.int32))
    pool.init(dev, len(fields))
    ctx.synchronize()
    t0 = time.perf_counter()
    steps = pool.run(t, fields, waves_per_cu=waves_per_cu)
    return steps, time.perf_counter() - t0, pool


# ---------------------------------------------------------------- curves of the sets

CHUNK_VERTICES = 134_217_727  # vertices per chunk of pa::integral_curves (particle_tracer.cpp:216-222)


def set_order(wl: BatchWorkload) -> np.ndarray:
    """Permutation of wl.particles that keeps each seed set contiguous, sets in
    set order, a set's seeds in locality order: the order a set's curves come
    out in (neighbouring lanes still gather neighbouring cells)."""
    sid = wl.particles["original_rank"].astype(np.int64)
    within = np.empty(len(sid), np.int64)
    loc = locality_order(wl.particles, wl.h)
    within[loc] = np.arange(len(loc))
    return np.lexsort((within, sid))


class _Arena:
    """First-fit allocator of slot ranges in the vertex arena."""

    def __init__(self, slots: int):
        self.free = [(0, int(slots))]  # (start, length), sorted by start

    def take(self, n: int) -> Optional[int]:
        for k, (s, l) in enumerate(self.free):
            if l >= n:
                self.free[k] = (s + n, l - n)
                if self.free[k][1] == 0:
                    del self.free[k]
                return s
        return None

    def give(self, s: int, n: int):
        self.free.append((s, n))
        self.free.sort()
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged


@dataclass
class CurveChunk:
    """One chunk of a finished set's curves, valid during the consumer call:
    device pointers and counts of its vertices (vec4), indices (int32, local
    to the chunk) and colours (vec4); seeds [seed_begin, seed_end) of the set."""
    set_id: int
    chunk: int
    seed_begin: int
    seed_end: int
    vertices: int
    n_vertices: int
    indices: int
    n_indices: int
    colors: int


def run_batch_vertices(ctx: Context, wl: BatchWorkload, fields: List[VectorField], waves_per_cu: int = 0,
                       arena_slots: Optional[int] = None, pass_iterations: int = 0, consumer=None,
                       color_mode: int = 2, color_parameter: float = 0.75, chunk_vertices: int = CHUNK_VERTICES,
                       order: Optional[np.ndarray] = None):
    """The batch with every set's curves, streamed out set by set.

    Sets are admitted in set order while their slots (size x R) fit the
    vertex arena (by default up to 30% of the free device memory); the pool
    advances every admitted particle by at most pass_iterations per pass (0:
    to its end; pa_gpu_pool_run_vertices: each particle writes its one-block
    reference slots into its set's part of the arena).  Bounded passes hand
    sets out sooner (on the batched workload: passes of 400 iterations 75
    ms of pool time, to the end 63 ms).  After each pass, a set whose particles have all ended is finished on the device
    in chunks of at most chunk_vertices slots, as the reference chunks its
    curves (chunk = chunk_vertices // R seeds): counted prune, index and colour
    in one pass over the slots (pa_gpu_finish_counted);
    each chunk goes to consumer(CurveChunk), then the set's arena range takes
    the next sets.  Returns totals (steps, vertices, indices, sets, passes,
    seconds)."""
    from . import lib, C  # noqa: F401  (ctypes entry points)
    order = set_order(wl) if order is None else order
    parts = wl.particles[order]
    sizes = np.array([s["size"] for s in wl.sets], np.int64)
    Rs = np.array([s["steps"] + 2 for s in wl.sets], np.int64)
    pool_off = np.concatenate([[0], np.cumsum(sizes)])
    sid = parts["original_rank"].astype(np.int64)
    if not np.array_equal(sid, np.repeat(np.arange(len(sizes)), sizes)):
        raise ValueError("run_batch_vertices: the particles must be set by set (set_order)")
    need = sizes * Rs
    if arena_slots is None:  # up to 30% of the free device memory
        free = ctx.device_info()["free_bytes"]
        arena_slots = int(max(need.max(), min(need.sum(), int(0.3 * free) // 16)))
    if need.max() > arena_slots:
        raise ValueError("a set needs %d slots, the arena holds %d" % (need.max(), arena_slots))
    if chunk_vertices < Rs.max():
        raise ValueError("chunk_vertices below a set's slots per seed")
    n = len(parts)
    t = make_tracer(ctx)
    dev = Particles.from_host(ctx, parts)
    pool = ParticlePool(ctx, n, requests=len(wl.sets))
    base = ctx.alloc(max(n, 1) * 8)
    counts = ctx.alloc(max(n, 1) * 4)
    arena = ctx.alloc(int(arena_slots) * 16)
    cap = int(min(chunk_vertices, max(need.max(), 1)))
    vbuf, ibuf, cbuf = ctx.alloc(cap * 16), ctx.alloc(cap * 4), ctx.alloc(cap * 16)
    offs = ctx.alloc((int(sizes.max()) + 1) * 8)
    alloc = _Arena(arena_slots)
    where = {}           # admitted set -> arena slot offset
    finished = set()
    next_set = 0
    n_hi = 0
    tot = dict(steps=0, vertices=0, indices=0, sets=0, passes=0, chunks=0)
    L = lib()
    t0 = time.perf_counter()

    def admit():
        nonlocal next_set, n_hi
        while next_set < len(sizes):
            s = next_set
            off = alloc.take(int(need[s]))
            if off is None:
                return
            where[s] = off
            a, b = int(pool_off[s]), int(pool_off[s + 1])
            base.upload(off + np.arange(b - a, dtype=np.int64) * Rs[s], byte_offset=a * 8)
            pc = pool._c(a)
            st = L.pa_gpu_pool_init(ctx.handle, C.c_void_p(dev.buf.at(48 * a)), C.c_size_t(b - a),
                                    C.c_int32(len(fields)), C.byref(pc))
            ctx.check(st, "pool_init")
            n_hi = max(n_hi, b)
            pool.n = max(pool.n, b)
            next_set += 1

    def finish(s):
        off, R = where.pop(s), int(Rs[s])
        a = int(pool_off[s])
        per = max(1, chunk_vertices // R)
        for k, j0 in enumerate(range(0, int(sizes[s]), per)):
            j1 = min(int(sizes[s]), j0 + per)
            m, ni = C.c_size_t(), C.c_size_t()
            ctx.check(L.pa_gpu_finish_counted(ctx.handle, C.c_void_p(arena.at((off + j0 * R) * 16)), C.c_size_t(R),
                                              C.c_void_p(counts.at((a + j0) * 4)), C.c_size_t(j1 - j0),
                                              C.c_void_p(vbuf.ptr), C.c_size_t(cap), C.byref(m),
                                              C.c_void_p(offs.ptr), C.c_void_p(ibuf.ptr), C.byref(ni),
                                              C.c_int(color_mode), C.c_float(color_parameter),
                                              C.c_void_p(cbuf.ptr)), "finish_counted")
            tot["vertices"] += m.value
            tot["indices"] += ni.value
            tot["chunks"] += 1
            if consumer is not None:
                consumer(CurveChunk(s, k, j0, j1, vbuf.ptr, m.value, ibuf.ptr, ni.value, cbuf.ptr))
        alloc.give(off, int(need[s]))
        finished.add(s)
        tot["sets"] += 1

    admit()
    while len(finished) < len(sizes):
        if not where:
            raise RuntimeError("run_batch_vertices: nothing admitted")
        tot["steps"] += pool.run_vertices(t, fields, base, arena, counts, n=n_hi, max_iterations=pass_iterations,
                                          waves_per_cu=waves_per_cu)
        tot["passes"] += 1
        done = pool.request_done.download(np.int32, len(sizes))
        for s in sorted(where):
            if done[s] == sizes[s]:
                finish(s)
        admit()
    ctx.synchronize()
    tot["seconds"] = time.perf_counter() - t0
    tot["arena_slots"] = int(arena_slots)
    return tot


def main(argv=None):
    """python -m pars_amd.batch: the batched workload on device 0 (BF16 cell
    records): `--runs` timed persistent launches, and with `--curves` the same
    batch with every set's curves at each of `--pass-iterations` (0: each pass
    to the end), one JSON line per run."""
    import argparse

    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--waves", type=int, default=0, help="waves per CU (0: as many as stay resident)")
    ap.add_argument("--curves", action="store_true")
    ap.add_argument("--pass-iterations", type=int, nargs="*", default=[0, 400])
    args = ap.parse_args(argv)
    ctx = Context(0)
    wl = make_workload()
    fields = build_fields(ctx, wl, BF16, LAYOUT_CELLS)
    steps, sec, pool = run_batch(ctx, wl, fields, args.waves)  # warm-up
    for r in range(args.runs):
        steps, sec, pool = run_batch(ctx, wl, fields, args.waves, pool=pool)
        ctx.synchronize()
        print(json.dumps(dict(run=r, particles=wl.n, steps=steps, seconds=sec, steps_per_s=steps / sec)), flush=True)
    if args.curves:
        for it in args.pass_iterations:
            ctx.timing_reset()
            ctx.timing(True)
            tot = run_batch_vertices(ctx, wl, fields, args.waves, pass_iterations=it)
            ctx.synchronize()
            ctx.timing(False)
            kt = ctx.kernel_time("pool_run")
            print(json.dumps(dict(pass_iterations=it, pool_run_ms=kt[0], **tot)), flush=True)
    for f in fields:
        f.release()


if __name__ == "__main__":
    main()
