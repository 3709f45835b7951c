"""Continuous-batching serving loop on the particle pool.

Requests (a seed set each, over fields kept in HBM) arrive over time.  The
loop admits every arrived request whose seeds fit into free pool slots
(first fit over slot ranges; the pool's capacity is set from free device
memory), then runs one bounded pass of the persistent pool kernel over all
admitted work, so requests that arrive while others are on the GPU join the
next pass and overlap with them on the device.  A request completes when
all of its particles are done (a per-request counter on the device); its
latency runs from arrival to the host seeing its endpoints.

The baseline processes the same arrivals one request at a time, each in its
own launch to completion (the synchronous model of pars' service loop,
pars/source/service.cpp:20-69).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import (BF16, FP32, Context, ParticlePool, Particles, VectorField, particle_tracer, PARTICLE_DTYPE)
from . import workloads as W
from .batch import locality_order

GRID = 1024
N_MODES = 64
DT = 0.002
STEPS = 500
REQ_SIZE = (256, 262_144)       # log-uniform seeds per request
BOX_FRACTION = (0.01, 0.10)     # of the domain volume


@dataclass
class Request:
    rid: int
    arrival: float                # seconds after the start
    particles: np.ndarray         # PARTICLE_DTYPE, original_rank set by the server
    admitted: float = math.nan
    done: float = math.nan
    endpoints: Optional[np.ndarray] = None
    steps: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def make_field(ctx: Context, storage=BF16, grid: int = GRID, n_modes: int = N_MODES, seed: int = W.DEFAULT_SEED):
    h = 2.0 * math.pi / (grid - 1)
    hf = np.float32(h)
    modes = W.fourier_modes(n_modes, seed)
    size = np.array([np.float32(grid) * hf] * 3, np.float32)
    return VectorField.synthesize(ctx, "fourier", modes, (grid, grid, grid), (0, 0, 0), h, np.zeros(3, np.float32),
                                  size, np.array([hf] * 3, np.float32), storage)


def make_requests(n: int, rate: float, grid: int = GRID, seed: int = W.DEFAULT_SEED, steps: int = STEPS,
                  size_range=REQ_SIZE, box_fraction=BOX_FRACTION) -> List[Request]:
    """Poisson arrivals at `rate` per second; each request has a log-uniform
    number of seeds, uniform in a random axis-aligned box of 1-10% of the
    domain volume, R = steps + 2, on field 0."""
    rng = W.SplitMix64(seed)
    h = np.float32(2.0 * math.pi / (grid - 1))
    extent = float(np.float32(grid - 1) * h)
    t = 0.0
    out = []
    lo, hi = math.log(size_range[0]), math.log(size_range[1])
    for r in range(n):
        t += -math.log(max(1.0 - rng.uniform(), 1e-300)) / rate
        size = int(round(math.exp(lo + (hi - lo) * rng.uniform())))
        frac = box_fraction[0] + (box_fraction[1] - box_fraction[0]) * rng.uniform()
        side = extent * frac ** (1.0 / 3.0)
        corner = np.array([rng.uniform() * (extent - side) for _ in range(3)])
        u = np.random.default_rng(rng.next_u64() & 0xFFFFFFFF).uniform(0.0, 1.0, (size, 3))
        p = np.zeros(size, PARTICLE_DTYPE)
        p["position"][:, :3] = (corner + u * side).astype(np.float32)
        p["remaining_iterations"] = steps + 2
        p["vector_field_index"] = 0
        p = p[locality_order(p, [float(h)])]  # neighbouring lanes trace neighbouring seeds
        out.append(Request(r, t, p))
    return out


class _Slots:
    """First-fit allocator of contiguous pool slot ranges."""

    def __init__(self, capacity: int):
        self.free = [(0, capacity)]

    def take(self, n: int) -> Optional[int]:
        for k, (a, b) in enumerate(self.free):
            if b - a >= n:
                self.free[k] = (a + n, b)
                if a + n == b:
                    del self.free[k]
                return a
        return None

    def give(self, a: int, n: int):
        self.free.append((a, a + n))
        self.free.sort()
        merged = []
        for x, y in self.free:
            if merged and merged[-1][1] == x:
                merged[-1] = (merged[-1][0], y)
            else:
                merged.append((x, y))
        self.free = merged

    def high_water(self, capacity: int) -> int:
        return self.free[-1][0] if self.free and self.free[-1][1] == capacity else capacity


def make_tracer(ctx: Context, dt: float = DT) -> particle_tracer:
    t = particle_tracer(ctx)
    t.set_integrator("runge_kutta_4")
    t.set_step_size(dt)
    return t


class Server:
    """Continuous batching over one pool: admit, run a bounded pass, retire."""

    def __init__(self, ctx: Context, fields: Sequence[VectorField], capacity: Optional[int] = None,
                 pass_iterations: int = 64, max_in_flight: int = 4096, memory_fraction: float = 0.25):
        self.ctx, self.fields = ctx, list(fields)
        if capacity is None:  # admission by free memory: a share of it, at 44 B of pool state per particle
            free = ctx.device_info()["free_bytes"]
            capacity = min(int(free * memory_fraction) // 44, 1 << 26)
        self.capacity = capacity
        self.pool = ParticlePool(ctx, capacity, requests=max_in_flight)
        self.tracer = make_tracer(ctx)
        self.pass_iterations = pass_iterations
        self.max_in_flight = max_in_flight

    def serve(self, requests: Sequence[Request], keep_endpoints: bool = True, clock=time.perf_counter):
        ctx, pool = self.ctx, self.pool
        slots = _Slots(self.capacity)
        ids = list(range(self.max_in_flight))[::-1]       # free request-counter slots
        waiting = list(requests)
        active = {}                                       # counter id -> (request, first slot)
        left = {}                                         # counter id -> iterations before it surely ends
        steps_total = 0
        t0 = clock()
        while waiting or active:
            now = clock() - t0
            # admit every arrived request that fits, in arrival order
            while waiting and waiting[0].arrival <= now and ids:
                req = waiting[0]
                n = len(req.particles)
                if n > self.capacity:
                    raise ValueError("request %d has more seeds than the pool holds" % req.rid)
                a = slots.take(n)
                if a is None:
                    break
                waiting.pop(0)
                cid = ids.pop()
                p = req.particles.copy()
                p["original_rank"] = cid
                pool.request_done.upload(np.zeros(1, np.int32), byte_offset=4 * cid)
                pool.init(Particles.from_host(ctx, p), len(self.fields), offset=a)
                req.admitted = clock() - t0
                active[cid] = (req, a)
                left[cid] = int(p["remaining_iterations"].max())  # iterations until its last particle ends
            if not active:
                if waiting:
                    wait = waiting[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.001))
                continue
            hw = slots.high_water(self.capacity)
            # pass length: short while requests wait (they join the next pass),
            # otherwise long enough for the soonest request to finish
            if waiting and waiting[0].arrival <= clock() - t0:
                k = self.pass_iterations
            else:
                k = max(self.pass_iterations, min(left[c] for c in active))
            steps_total += pool.run(self.tracer, self.fields, n=hw, max_iterations=k)
            for c in active:
                left[c] -= k
            done = pool.request_done.download(np.int32, self.max_in_flight)
            for cid in [c for c, (r, a) in active.items() if done[c] == len(r.particles)]:
                req, a = active.pop(cid)
                n = len(req.particles)
                if keep_endpoints:
                    req.endpoints = pool.position.download(np.float32, 4 * n, byte_offset=16 * a).reshape(n, 4)
                req.steps = int(pool.steps.download(np.int32, n, byte_offset=4 * a).sum())
                req.done = clock() - t0
                slots.give(a, n)
                ids.append(cid)
        return steps_total, clock() - t0


class ResidentServer:
    """Serving on the device-resident loop (pa_gpu_serve_*): a request is published as soon as it
    arrives and fits the ring; its latency runs from arrival to the host holding its endpoints."""

    def __init__(self, ctx: Context, fields: Sequence[VectorField], capacity: Optional[int] = None,
                 request_slots: int = 4096, waves_per_cu: int = 0, memory_fraction: float = 0.02):
        from . import DeviceServer
        self.ctx, self.fields = ctx, list(fields)
        if capacity is None:  # admission by memory: 68 B of pinned ring per particle (in + out), 20 B on device
            free = ctx.device_info()["free_bytes"]
            capacity = min(int(free * memory_fraction) // 68, 1 << 24)
        self.capacity = capacity
        self.tracer = make_tracer(ctx)
        self.server = DeviceServer(ctx, self.tracer, self.fields, capacity, request_slots, waves_per_cu)

    def serve(self, requests: Sequence[Request], keep_endpoints: bool = True, clock=time.perf_counter,
              timeout: float = 60.0):
        waiting = list(requests)
        active = {}  # slot -> request
        steps_total = 0
        t0 = clock()
        while waiting or active:
            now = clock() - t0
            if now > timeout + (requests[-1].arrival if requests else 0.0):
                raise RuntimeError("resident serving: %d requests still open after %.0f s" % (len(active), now))
            while waiting and waiting[0].arrival <= now:
                req = waiting[0]
                if len(req.particles) > self.capacity:
                    raise ValueError("request %d has more seeds than the ring holds" % req.rid)
                slot = self.server.submit(req.particles)
                if slot is None:
                    break
                waiting.pop(0)
                req.admitted = clock() - t0
                active[slot] = req
            for slot in [s for s in active if self.server.done(s)]:
                req = active.pop(slot)
                ends, steps, total = self.server.collect(slot, len(req.particles), keep_endpoints)
                req.endpoints = ends
                req.steps = total
                req.done = clock() - t0
                steps_total += total
            if not active and waiting:
                wait = waiting[0].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.0005))
        wall = clock() - t0
        self.server.stop()  # other work of the process may share the kernel's hardware queue
        return steps_total, wall

    def close(self):
        self.server.stop()
        self.server.close()


def serve_one_at_a_time(ctx: Context, fields: Sequence[VectorField], requests: Sequence[Request],
                        keep_endpoints: bool = True, clock=time.perf_counter):
    """Baseline: each request alone, in arrival order, one launch to
    completion (device buffers allocated once, for the largest request)."""
    t = make_tracer(ctx)
    pool = ParticlePool(ctx, max(len(r.particles) for r in requests))
    steps_total = 0
    t0 = clock()
    for req in requests:
        now = clock() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        req.admitted = clock() - t0
        n = len(req.particles)
        pool.init(Particles.from_host(ctx, req.particles), len(fields))
        steps_total += pool.run(t, fields, n=n)
        if keep_endpoints:
            req.endpoints = pool.position.download(np.float32, 4 * n).reshape(n, 4)
        req.steps = int(pool.steps.download(np.int32, n).sum())
        req.done = clock() - t0
    return steps_total, clock() - t0


def latency_stats(requests: Sequence[Request]) -> dict:
    lat = np.array([r.latency for r in requests]) * 1e3
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()))
