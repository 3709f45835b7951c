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
             waves=(4, 8, 16, 32), sample_stride: int = 16, cache: Optional[str] = None) -> Tuning:
    """Time every (layout, waves per CU) on every sample_stride-th seed of the
    batch and return the fastest.  With `cache`, a previous result for the
    same device and batch shape is reused."""
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
            pool.init(dev, len(fields))
            ctx.synchronize()
            t0 = time.perf_counter()
            steps = pool.run(t, fields, waves_per_cu=wv)
            dt = time.perf_counter() - t0
            rate = steps / dt
            table["%s/w%d" % ("cells" if lay == LAYOUT_CELLS else "nodes", wv)] = rate
            if best is None or rate > best[2]:
                best = (lay, wv, rate)
        for f in fields:
            f.release()
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
