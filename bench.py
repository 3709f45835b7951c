#!/usr/bin/env python3
"""Benchmark of the per-process advection path on one MI355X.

Workload (one rank's share of a 512^3 ABC domain split 4x4x4): the own
block of rank 21 and its six face neighbours, 7 x 129^3 ghosted blocks
resident in HBM; 262,144 seeds on a stride-2 lattice over the own block;
runge_kutta_4 with the reference's frozen-derivative stepper, dt = 0.005,
up to 1000 steps (remaining_iterations = 1002); particles continue across
the 7 blocks and stop on leaving their union or when |u| < 1e-6.

One step = one pass of the hot path over that batch: seed lattice ->
union trace -> prune -> index -> colour, all on the device with the inputs
(fields) already resident.  The value is integrator steps (particle-steps)
per second of that whole pass, FP32 field, with passes queued back to back
the way pars' round loop runs rounds (pipeline.cpp:153-209): each pass's
trace follows the previous trace, and its finish (prune, index, colour)
runs beside the next pass's trace (the pipelined fused round).  The
synchronous form (one pass at a time, a wait after each) gives the pass
latency and the trace kernel's roofline.  BF16 / FP16 field storage and
the consistent-weight interpolation are reported beside it, with their
endpoint error against the FP64 CPU oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 particle-steps/s on 1 MI355X (FP32 vs BF16 field) + % HBM peak; p99 latency"
HBM_PEAK_GBS = 8000.0
R = 1002          # remaining_iterations: 1000 integrator steps
DT = 0.005
EXTRA_SLOTS = 30  # per-seed region beyond R for union segments
GATHER_BYTES = {0: 96, 1: 48, 2: 48}  # 8 corners x 3 components x storage bytes
STORE_BYTES = 16                      # one vec4 FP32 vertex per step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Pipeline:
    """Device buffers and calls of one step, reused across steps."""

    def __init__(self, pa, ctx, fields, share, storage, interp, n_max, order="step", engine=None, fused=False,
                 graph=False, pipelined=False):
        self.pa, self.ctx, self.lib = pa, ctx, pa.lib()
        self.order = order
        self.fused = fused
        self.graph = None
        self.pipelined = pipelined
        self.share = share
        t = pa.particle_tracer(ctx)
        t.set_local_vector_field(fields[0])
        t.set_neighbor_vector_fields(fields[1:], share.neighbor_ranks)
        t.set_neighbor_face_masks(share.face_masks)
        t.set_integrator(pa.RUNGE_KUTTA_4)
        t.set_step_size(DT)
        t.set_stop_rule(pa.STOP_NORM, 1e-6)
        t.set_interpolation(interp)
        if order == "chunked":
            t.set_engine(pa.ENGINE_QUAD)
        elif engine is not None:
            t.set_engine(engine)
        self.tracer = t
        self.region = R + EXTRA_SLOTS
        if order == "chunked":
            self.region = (self.region + 7) // 8 * 8
        self.seeds = ctx.alloc(n_max * 48)
        self.ub = t.union_buffers(n_max, self.region, endpoints=True)
        self.cap = n_max * self.region
        self.pruned = ctx.alloc(self.cap * 16)
        self.offs = ctx.alloc((n_max + 1) * 8)
        self.indices = ctx.alloc(self.cap * 4)
        self.colors = ctx.alloc(self.cap * 16)
        loc = share.local
        self.seed_args = (pa._f3(loc.offset), pa._f3(loc.size), pa._f3(loc.spacing * np.float32(2)))
        if graph:
            self._capture(n_max)
        if pipelined:
            # two sets of round buffers, alternated: a round finishes while the next one traces
            self.sets = [dict(seeds=self.seeds, ub=self.ub, pruned=self.pruned, indices=self.indices,
                              colors=self.colors, result=ctx.alloc(48))]
            ub = t.union_buffers(n_max, self.region, endpoints=True)
            self.sets.append(dict(seeds=ctx.alloc(n_max * 48), ub=ub, pruned=ctx.alloc(self.cap * 16),
                                  indices=ctx.alloc(self.cap * 4), colors=ctx.alloc(self.cap * 16),
                                  result=ctx.alloc(48)))
            self.k = 0
            self.host_result = np.zeros(1, pa.ROUND_RESULT_DTYPE)

    def _step_pipelined(self):
        """Queue one pass on the next set of buffers (seed lattice, then the pipelined round: its
        finish runs on the context's second stream under the next pass's trace); no read-back."""
        pa, lib, ctx = self.pa, self.lib, self.ctx
        b = self.sets[self.k % 2]
        self.k += 1
        n = C.c_size_t()
        o, s, st = self.seed_args
        ctx.check(lib.pa_gpu_seed_generate(ctx.handle, o, s, st, C.c_int32(R), C.c_int32(21),
                                           C.c_void_p(b["seeds"].ptr), C.c_size_t(b["seeds"].nbytes // 48),
                                           C.byref(n)), "seeds")
        bufs = dict(b["ub"], out_vertices=b["pruned"], out_indices=b["indices"], out_colors=b["colors"])
        self.tracer.trace_union_round_pipelined(pa.Particles(b["seeds"], n.value), self.region, bufs, b["result"],
                                                2, 0.75, vertex_capacity=self.cap)
        self.n_pipe = n.value
        # what the error and parity legs read: the buffers of this pass
        self.seeds, self.ub, self.pruned = b["seeds"], b["ub"], b["pruned"]
        self.indices, self.colors = b["indices"], b["colors"]

    def collect(self):
        """Wait for every queued pass; the last pass's counts (pipelined form)."""
        pa, lib, ctx = self.pa, self.lib, self.ctx
        ctx.synchronize()
        b = self.sets[(self.k - 1) % 2]
        ctx.check(lib.pa_gpu_memcpy_d2h(ctx.handle, self.host_result.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(b["result"].ptr), C.c_size_t(48)), "result")
        r = self.host_result
        ctx.check(lib.pa_gpu_round_result_check(ctx.handle, r.ctypes.data_as(C.c_void_p), C.c_size_t(self.cap),
                                                C.c_size_t(self.cap)), "round")
        if r["n_continuation"][0]:
            raise RuntimeError("segment region overflow")
        return dict(particles=self.n_pipe, steps=int(r["total_steps"][0]), vertices=int(r["n_vertices"][0]),
                    indices=int(r["n_indices"][0]))

    def _capture(self, n_max):
        """One pass (seed lattice + asynchronous fused round) captured as a graph; each step
        replays it and reads the 48-byte result record back."""
        pa, lib, ctx = self.pa, self.lib, self.ctx
        b = dict(self.ub)
        b.update(out_vertices=self.pruned, out_indices=self.indices, out_colors=self.colors)
        self.result = ctx.alloc(48)
        parts = pa.Particles(self.seeds, n_max)
        o, s, st = self.seed_args
        n = C.c_size_t()
        ctx.check(lib.pa_gpu_seed_generate(ctx.handle, o, s, st, C.c_int32(R), C.c_int32(21),
                                           C.c_void_p(self.seeds.ptr), C.c_size_t(n_max), C.byref(n)), "seeds")
        self.tracer.trace_union_round_async(parts, self.region, b, self.result, 2, 0.75)  # sizes the workspace
        ctx.synchronize()
        with pa.Graph.capture(ctx) as g:
            ctx.check(lib.pa_gpu_seed_generate(ctx.handle, o, s, st, C.c_int32(R), C.c_int32(21),
                                               C.c_void_p(self.seeds.ptr), C.c_size_t(n_max), C.byref(n)), "seeds")
            self.tracer.trace_union_round_async(parts, self.region, b, self.result, 2, 0.75)
        self.graph = g
        self.n_graph = n.value
        self.host_result = np.zeros(1, pa.ROUND_RESULT_DTYPE)

    def step(self):
        pa, lib, ctx = self.pa, self.lib, self.ctx
        if self.pipelined:
            self._step_pipelined()
            return self.collect()
        if self.graph is not None:
            self.graph.launch()
            ctx.check(lib.pa_gpu_memcpy_d2h(ctx.handle, self.host_result.ctypes.data_as(C.c_void_p),
                                            C.c_void_p(self.result.ptr), C.c_size_t(48)), "result")
            r = self.host_result
            ctx.check(lib.pa_gpu_round_result_check(ctx.handle, r.ctypes.data_as(C.c_void_p), C.c_size_t(self.cap),
                                                    C.c_size_t(self.cap)), "round")
            if r["n_continuation"][0]:
                raise RuntimeError("segment region overflow")
            return dict(particles=self.n_graph, steps=int(r["total_steps"][0]), vertices=int(r["n_vertices"][0]),
                        indices=int(r["n_indices"][0]))
        n = C.c_size_t()
        o, s, st = self.seed_args
        ctx.check(lib.pa_gpu_seed_generate(ctx.handle, o, s, st, C.c_int32(R), C.c_int32(21),
                                           C.c_void_p(self.seeds.ptr), C.c_size_t(self.seeds.nbytes // 48),
                                           C.byref(n)), "seeds")
        parts = pa.Particles(self.seeds, n.value)
        if self.fused:
            ncont, nv, ni, steps = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
            ctx.check(lib.pa_gpu_tracer_trace_union_round(
                self.tracer._h, C.c_void_p(self.seeds.ptr), n, C.c_size_t(self.region),
                C.c_void_p(self.ub["vertices"].ptr), C.c_void_p(self.ub["counts"].ptr),
                C.c_void_p(self.ub["cont"].ptr), C.byref(ncont), C.c_int(2), C.c_float(0.75),
                C.c_void_p(self.pruned.ptr), C.c_size_t(self.cap), C.byref(nv), C.c_void_p(self.indices.ptr),
                C.c_size_t(self.cap), C.byref(ni), C.c_void_p(self.colors.ptr), C.c_void_p(self.ub["endpoints"].ptr),
                C.c_void_p(self.ub["seed_steps"].ptr), C.byref(steps)), "round")
            if ncont.value:
                raise RuntimeError("segment region overflow: %d seeds need a continuation pass" % ncont.value)
            return dict(particles=n.value, steps=int(steps.value), vertices=nv.value, indices=ni.value)
        r = self.tracer.trace_union(parts, self.region, buffers=self.ub, order=self.order)
        if r["continuation"].n:
            raise RuntimeError("segment region overflow: %d seeds need a continuation pass" % r["continuation"].n)
        m = C.c_size_t()
        if self.order == "group8":
            ctx.check(lib.pa_gpu_prune_counted_group8(
                ctx.handle, C.c_void_p(self.ub["vertices"].ptr), C.c_size_t(self.region),
                C.c_void_p(self.ub["counts"].ptr), n, C.c_void_p(self.pruned.ptr), C.c_size_t(self.cap),
                C.byref(m), C.c_void_p(self.offs.ptr)), "prune")
        elif self.order == "chunked":
            ctx.check(lib.pa_gpu_prune_counted_chunked(
                ctx.handle, C.c_void_p(self.ub["vertices"].ptr), C.c_size_t(self.region),
                C.c_void_p(self.ub["counts"].ptr), n, C.c_void_p(self.pruned.ptr), C.c_size_t(self.cap),
                C.byref(m), C.c_void_p(self.offs.ptr)), "prune")
        elif self.order == "step":
            ctx.check(lib.pa_gpu_prune_counted_step_major(
                ctx.handle, C.c_void_p(self.ub["vertices"].ptr), n, C.c_void_p(self.ub["counts"].ptr),
                C.c_void_p(self.pruned.ptr), C.c_size_t(self.cap), C.byref(m), C.c_void_p(self.offs.ptr)), "prune")
        else:
            ctx.check(lib.pa_gpu_prune_counted(
                ctx.handle, C.c_void_p(self.ub["vertices"].ptr), C.c_size_t(self.region),
                C.c_void_p(self.ub["counts"].ptr), n, C.c_void_p(self.pruned.ptr), C.c_size_t(self.cap),
                C.byref(m), C.c_void_p(self.offs.ptr)), "prune")
        ni = C.c_size_t()
        ctx.check(lib.pa_gpu_index_generate(ctx.handle, C.c_void_p(self.pruned.ptr), m, C.c_void_p(self.indices.ptr),
                                            C.byref(ni)), "index")
        ctx.check(lib.pa_gpu_color_generate(ctx.handle, C.c_void_p(self.pruned.ptr), m, C.c_void_p(self.indices.ptr),
                                            ni, C.c_int(2), C.c_float(0.75), C.c_void_p(self.colors.ptr)), "color")
        return dict(particles=n.value, steps=r["total_steps"], vertices=m.value, indices=ni.value)


def tune_pipeline(pa, ctx, p):
    """The fused round's launch constants for this workload (pars_amd.tune), set on the pipeline's tracer."""
    from pars_amd import tune as T
    n = C.c_size_t()
    o, s, st = p.seed_args
    ctx.check(pa.lib().pa_gpu_seed_generate(ctx.handle, o, s, st, C.c_int32(R), C.c_int32(21),
                                            C.c_void_p(p.seeds.ptr), C.c_size_t(p.seeds.nbytes // 48),
                                            C.byref(n)), "seeds")
    b = dict(p.ub, out_vertices=p.pruned, out_indices=p.indices, out_colors=p.colors)
    return T.tune_round(ctx, p.tracer, pa.Particles(p.seeds, n.value), p.region, b, ctx.alloc(48))


def run_form(pa, ctx, fields, share, storage, interp, n_seeds, steps, warmup, order, engine=None, fused=False,
             graph=False, pipelined=False, tune=False, launch=None):
    p = Pipeline(pa, ctx, fields, share, storage, interp, n_seeds, order, engine, fused, graph, pipelined)
    tuning = None
    if fused and launch is not None:  # given constants (e.g. for a profiled run), no tuning
        p.tracer.set_launch(launch[0], launch[1])
        p.tracer.set_xcd_remap(int(launch[2]))
        tuning = dict(quad_switch=launch[0], pair_live=launch[1], xcd_remap=launch[2], given=True)
    elif tune and fused and not graph:
        tuning = tune_pipeline(pa, ctx, p).as_dict()
    for _ in range(warmup):
        info = p.step()
    ctx.synchronize()
    ctx.timing_reset()
    ctx.timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    total = 0
    lat = []
    if pipelined:
        # passes queued back to back: each pass's trace waits for the previous trace, its finish
        # overlaps the next trace; one wait at the end (no per-pass latency in this form)
        for _ in range(steps):
            p._step_pipelined()
        info = p.collect()
        total = info["steps"] * steps
        lat = [float("nan")]
    else:
        for _ in range(steps):
            t1 = time.perf_counter()
            info = p.step()
            ctx.synchronize()
            lat.append(time.perf_counter() - t1)
            total += info["steps"]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.timing(False)
    kt = {k: ctx.kernel_time(k) for k in ("seeds", "trace_union", "prune_scan", "prune_counted", "index", "color",
                                          "round_counts", "round_finish")}
    trace_ms = kt["trace_union"][0] / kt["trace_union"][1] if kt["trace_union"][1] else None  # None: graph replays
    return p, dict(wall_s=wall, ms_per_step=1e3 * wall / steps, steps_per_pass=total // steps,
                   tuning=tuning,
                   pass_latency_ms=dict(p50=1e3 * float(np.percentile(lat, 50)),
                                        p99=1e3 * float(np.percentile(lat, 99)), max=1e3 * max(lat)),
                   value=total / wall, trace_ms=trace_ms, info=info,
                   stage_ms={k: v[0] / max(v[1], 1) for k, v in kt.items()})


def endpoint_errors(pa, orc, p, share, storage, interp, sample_stride):
    """Endpoints of a sample of seeds against the FP64 oracle (consistent
    interpolant, same stepper, same union walk and stop rule)."""
    from pars_amd import workloads as W

    ctx = p.ctx
    p.step()
    n = p.seeds.nbytes // 48
    seeds = p.seeds.download(pa.PARTICLE_DTYPE, n)
    ends = p.ub["endpoints"].download(np.float32, 4 * n).reshape(n, 4)
    ssteps = p.ub["seed_steps"].download(np.int32, n)
    sel = np.arange(0, n, sample_stride)
    blocks = [share.local] + list(share.neighbors)
    f64 = [orc.Field64(W.abc_block_f64(b, share.h), b.offset.astype(np.float64), b.size.astype(np.float64),
                       [share.h] * 3) for b in blocks]
    e64, s64 = orc.endpoints_f64(seeds["position"][sel].astype(np.float64), R, f64[0], neighbors=f64[1:],
                                 neighbor_rank=share.neighbor_ranks, face_masks=share.face_masks, dt=DT,
                                 integrator=orc.RK4, union=True, stop_rule=orc.STOP_NORM, stop_eps=1e-6)
    same = ssteps[sel] == s64
    d = np.linalg.norm(ends[sel][same, :3].astype(np.float64) - e64[same, :3], axis=1)
    extent = float(share.local.size[0])
    return dict(sample=int(len(sel)), same_step_count_frac=float(same.mean()),
                max_endpoint_err=float(d.max()) if len(d) else None,
                rms_endpoint_err=float(np.sqrt((d ** 2).mean())) if len(d) else None,
                max_err_over_block_extent=float(d.max() / extent) if len(d) else None,
                p99_endpoint_err=float(np.percentile(d, 99)) if len(d) else None,
                _ends=ends[sel].copy(), _steps=ssteps[sel].copy())


def storage_error(e16, e32, extent):
    """A 16-bit form's sampled endpoints against its FP32 twin on the same
    interpolant (seeds with the same step count): the storage part of the
    error, apart from the FP32 arithmetic floor against FP64."""
    same = e16["_steps"] == e32["_steps"]
    d = np.linalg.norm(e16["_ends"][same, :3].astype(np.float64) - e32["_ends"][same, :3].astype(np.float64),
                       axis=1)
    return dict(sample=int(len(same)), same_step_count_frac=float(same.mean()),
                max_endpoint_err=float(d.max()) if len(d) else None,
                rms_endpoint_err=float(np.sqrt((d ** 2).mean())) if len(d) else None,
                max_err_over_block_extent=float(d.max() / extent) if len(d) else None,
                p99_endpoint_err=float(np.percentile(d, 99)) if len(d) else None)


def host_info(threads):
    """Host CPU of this run: model, logical CPUs the OS reports and may
    schedule this process on, and the threads the baseline used."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        import subprocess
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except Exception:  # reported as unknown
        nproc = None
    return dict(cpu_model=model, nproc=nproc, os_cpu_count=os.cpu_count(),
                affinity=len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
                hardware_concurrency=None, omp_num_threads=os.environ.get("OMP_NUM_THREADS"), threads_used=threads)


def cpu_baseline(orc, share, oracle_fields, threads, gpu=None, min_seconds=10.0):
    """The oracle's FP32 restatement of the same union trace (all 262,144
    seeds, vertex output) on host cores, over the node values the device
    holds (read back from the FP32 fields).  Its first pass is also the
    parity check of the headline pass: per-seed slot counts, step count and
    every vertex, byte for byte."""
    seeds = orc.seed_generate(share.local.offset, share.local.size, share.local.spacing * np.float32(2), R, 21)
    orc.set_threads(threads)
    steps, passes = 0, 0
    parity = None
    t0 = time.perf_counter()
    while passes == 0 or time.perf_counter() - t0 < min_seconds:
        u = orc.trace_union(seeds, oracle_fields[0], oracle_fields[1:], share.neighbor_ranks, share.face_masks,
                            dt=DT, integrator=orc.RK4, stop_rule=orc.STOP_NORM, stop_eps=1e-6,
                            vertex_capacity=len(seeds) * (R + EXTRA_SLOTS))
        if passes == 0 and gpu is not None:
            t_check = time.perf_counter()
            cnt = np.bincount(u["segment_seed"], weights=u["segment_length"], minlength=len(seeds)).astype(np.int64)
            parity = dict(form=gpu["form"], seeds=len(seeds), steps_equal=bool(u["total_steps"] == gpu["steps"]),
                          counts_equal=bool(np.array_equal(cnt, gpu["counts"].astype(np.int64))),
                          vertices=int(len(u["vertices"])),
                          vertices_bitexact=bool(len(u["vertices"]) == len(gpu["vertices"]) and
                                                 u["vertices"].tobytes() == gpu["vertices"].tobytes()),
                          seeds_bitexact=bool(seeds.tobytes() == gpu["seeds"].tobytes()))
            parity["ok"] = all(parity[k] for k in ("steps_equal", "counts_equal", "vertices_bitexact",
                                                   "seeds_bitexact"))
            t0 += time.perf_counter() - t_check  # the comparison is not CPU-baseline time
        steps += u["total_steps"]
        passes += 1
        del u
    dt = time.perf_counter() - t0
    hi = host_info(threads)
    hi["hardware_concurrency"] = orc.hardware_concurrency()
    return dict(value=steps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d passes of the union trace with vertex output over all %d seeds (%d steps per pass), "
                       "%.1f s, %d threads" % (passes, len(seeds), steps // passes, dt, threads),
                host=hi), parity


def kernel_name(form):
    tokens = form.split("-")
    st = tokens[0].upper()
    lay = "cells" if "cells" in tokens else "nodes"
    interp = "consistent" if "consistent" in tokens else "reference"
    if "lane" in tokens:
        k = "k_trace_union"
    elif "quad" in tokens:
        k = "k_quad_union"
    else:
        k = "k_hybrid_union"
    return "%s<%s %s, RK4, %s>" % (k, st, lay, interp)


def bf16_ratio(forms, extra):
    """BF16 over FP32 with the headline form's layout and engine."""
    head = forms[0]
    twin = "bf16" + head[len("fp32"):] if head.startswith("fp32") else None
    if twin not in extra:
        return None
    return dict(form=twin, pass_ratio=extra[twin]["value"] / extra[head]["value"],
                trace_ratio=extra[twin]["trace_steps_per_s"] / extra[head]["trace_steps_per_s"])


FTLE_FORMS = (("fp32-cells", "FP32", "cells"), ("fp32", "FP32", "nodes"), ("bf16-cells", "BF16", "cells"),
              ("fp16-cells", "FP16", "cells"), ("bf16", "BF16", "nodes"), ("fp32-cells-consistent", "FP32", "cells"))


def bench_ftle(pa, ctx, orc, n=256, steps=1000, dt=0.01, forms=FTLE_FORMS):
    """FTLE workload: ABC on n^3 nodes (one block), one seed per node, `steps`
    RK4 steps (advector R = steps + 1), T = steps * dt; flow map in FP32
    arithmetic over the field stored FP32 / BF16 / FP16 (node bricks or cell
    records), Jacobian / Cauchy-Green / eigen in FP64.  Parity at full size
    (every form, the oracle reading the same stored values): sampled
    flow-map voxels against the oracle's advector (bit-exact), and the FTLE
    of an interior sub-cube against the oracle's FP64 post-processing of the
    same flow-map values (1e-4 relative).  Error: the FTLE of that sub-cube
    against an FP64 flow map of the analytic field (the oracle's FP64 twin,
    consistent interpolation), relative to max(|FTLE|, 1e-3); with seeds on
    the nodes, the reference's interpolation defect (a quotient that rounds
    up to the node reads the next node) dominates it for every storage form
    with the reference weights; the consistent-weight FP32 form shows the
    FP32 arithmetic floor alone.  The storage error alone is the FTLE of a
    16-bit form against the FP32 form's on the same sub-cube."""
    from pars_amd import workloads as W

    h = 2.0 * np.pi / (n - 1)
    hf = np.float32(h)
    size = np.array([np.float32(n) * hf] * 3, np.float32)
    sp = np.array([hf] * 3, np.float32)
    T = steps * dt
    c0, m = n // 2 - 16, 34  # interior sub-cube for the FTLE checks
    truth = None
    if orc is not None:
        # FP64 flow map of the sub-cube's seeds (tracer semantics: R = steps + 2
        # gives `steps` steps and the same freeze on leaving the block)
        blk = W.BlockInfo(0, (0, 0, 0), (0, 0, 0), (n, n, n))
        f64 = orc.Field64(W.abc_block_f64(blk, h), np.zeros(3), size.astype(np.float64), [h] * 3)
        ijk = np.stack(np.meshgrid(*[np.arange(c0, c0 + m)] * 3, indexing="ij"), -1).reshape(-1, 3)
        pos = np.zeros((len(ijk), 4), np.float64)  # x, y, z, w per seed
        pos[:, :3] = (hf * ijk.astype(np.float32)).astype(np.float64)
        e64, _ = orc.endpoints_f64(pos, steps + 2, f64, dt=dt, integrator=orc.RK4)
        truth, _ = orc.ftle(e64[:, :3].reshape(m, m, m, 3).astype(np.float32), sp, T)
        truth = truth[1:-1, 1:-1, 1:-1].astype(np.float64)
        del f64
    out_forms = {}
    fp32_sub = None  # FTLE sub-cube of the first FP32 form (storage error of the others)
    for name, st, lay in forms:
        field = pa.VectorField.synthesize(ctx, "abc", W.ABC, (n, n, n), (0, 0, 0), h, np.zeros(3, np.float32), size,
                                          sp, getattr(pa, st))
        if lay == "cells":
            field.set_layout(pa.LAYOUT_CELLS)
        fm = pa.flow_map_generator(ctx)
        fm.set_vector_field(field)
        consistent = "consistent" in name.split("-")
        interp = pa.INTERP_CONSISTENT if consistent else pa.INTERP_REFERENCE
        fm.tracer.set_interpolation(interp)
        fm.set_integrator(pa.RUNGE_KUTTA_4)
        fm.set_step_size(dt)
        out, dims, spacing, _ = fm.generate((n, n, n), steps + 1, 1.0)  # warm-up
        ftle = pa.ftle_map_generator(ctx)
        ftle.set_flow_map(out, dims, spacing)
        ftle.set_time(T)
        ftle.generate()  # warm-up
        ctx.synchronize()
        ctx.timing_reset()
        ctx.timing(True)
        t0 = time.perf_counter()
        out, dims, spacing, nsteps = fm.generate((n, n, n), steps + 1, 1.0)
        ftle.set_flow_map(out, dims, spacing)
        ftle.set_time(T)
        fv, _ = ftle.generate()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        ctx.timing(False)
        fms = ctx.kernel_time("flow_map")
        fts = ctx.kernel_time("ftle")
        flow_ms = fms[0] / max(fms[1], 1)
        ftle_ms = fts[0] / max(fts[1], 1)
        bytes_per_step = 96 if st == "FP32" else 48
        res = dict(workload="ABC %d^3 %s %s, %d seeds (one per node), %d RK4 steps, dt %g, T %g; FP32 arithmetic, "
                            "FP64 FTLE" % (n, st, lay, n ** 3, steps, dt, T),
                   steps=nsteps, flow_map_ms=flow_ms, ftle_ms=ftle_ms, wall_ms=1e3 * wall,
                   flow_map_steps_per_s=nsteps / (flow_ms * 1e-3), ftle_nodes_per_s=n ** 3 / (ftle_ms * 1e-3),
                   roofline=dict(bound="hbm", algorithmic_bytes_per_step=bytes_per_step,
                                 achieved_gbs=nsteps * bytes_per_step / (flow_ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS))
        if orc is not None:
            flow = out.download(np.float32, n ** 3 * 3).reshape(n, n, n, 3)
            gsub = fv.download(np.float32, n ** 3).reshape(n, n, n)[c0 + 1:c0 + m - 1, c0 + 1:c0 + m - 1,
                                                                    c0 + 1:c0 + m - 1].astype(np.float64)
            # flow-map parity on sampled voxels (oracle on the stored values)
            rng = np.random.default_rng(5)
            ij = rng.integers(0, n, (512, 3))
            p = np.zeros(len(ij), orc.PARTICLE_DTYPE)
            p["position"][:, :3] = (hf * ij.astype(np.float32)).astype(np.float32)
            p["remaining_iterations"] = steps + 1
            p["vector_field_index"] = -1
            of = orc.Field(field.read(), np.zeros(3, np.float32), size, sp)
            q, _, _, _ = orc.advect(p, of, dt=dt, integrator=orc.RK4,
                                    interp=orc.INTERP_CONSISTENT if consistent else orc.INTERP_REFERENCE)
            got = flow[ij[:, 0], ij[:, 1], ij[:, 2]]
            res["flow_map_sample_bitexact"] = bool(got.tobytes() == q["position"][:, :3].copy().tobytes())
            # FP64 post-processing of the same flow-map values
            sub = flow[c0:c0 + m, c0:c0 + m, c0:c0 + m]
            of_, _ = orc.ftle(sub, sp, T)
            want = of_[1:-1, 1:-1, 1:-1]
            res["ftle_max_rel_err_vs_oracle"] = float(
                (np.abs(gsub - want) / np.maximum(np.abs(want), 1e-3)).max())
            rel = np.abs(gsub - truth) / np.maximum(np.abs(truth), 1e-3)
            res["ftle_err_vs_fp64"] = dict(max_rel=float(rel.max()), rms_rel=float(np.sqrt((rel ** 2).mean())),
                                           p99_rel=float(np.percentile(rel, 99)), nodes=int(rel.size))
            if fp32_sub is None and st == "FP32" and not consistent:
                fp32_sub = gsub
            elif fp32_sub is not None:
                rel = np.abs(gsub - fp32_sub) / np.maximum(np.abs(fp32_sub), 1e-3)
                # a 16-bit form against the FP32 one: the storage error; the consistent form against it:
                # what the interpolation defect adds
                key = "ftle_diff_vs_fp32_reference_interp" if consistent else "ftle_err_vs_fp32_form"
                res[key] = dict(max_rel=float(rel.max()), rms_rel=float(np.sqrt((rel ** 2).mean())),
                                p99_rel=float(np.percentile(rel, 99)))
        field.release()
        out_forms[name] = res
    head = dict(out_forms[forms[0][0]])
    head["forms"] = out_forms
    return head


def scattered_line_rate(pa, ctx, nbytes, line=64, read=48, n_reads=1 << 30):
    """Reads per second of pa_gpu_gather_lines: read-byte reads at the start of
    pseudo-random line-byte lines of an nbytes buffer, 8 in flight per lane."""
    lib = pa.lib()
    src, out = ctx.alloc(nbytes), ctx.alloc(16)

    def run(n, seed):
        ctx.check(lib.pa_gpu_gather_lines(ctx.handle, C.c_void_p(src.ptr), C.c_size_t(nbytes), C.c_int32(line),
                                          C.c_int32(read), C.c_int64(n), C.c_uint64(seed), C.c_void_p(out.ptr)),
                  "gather_lines")

    run(n_reads // 8, 1)  # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    run(n_reads, 2)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    src.free()
    out.free()
    return n_reads / sec


def bench_batched(pa, ctx, cache=None):
    """Batched uneven work: 8 random-Fourier curl fields (64^3 .. 256^3, BF16),
    64 seed sets (log-uniform 1e3..1e6 seeds, 100..2000 steps, dt 0.01) in one
    persistent work-queue kernel, launch shape from the autotuner; the round
    (pool initialisation + run) captured as a graph and replayed."""
    from pars_amd import batch as B

    wl = B.make_workload()
    tun = B.autotune(ctx, wl, cache=cache)
    fields = B.build_fields(ctx, wl, pa.BF16, tun.layout)
    steps, sec, pool = B.run_batch(ctx, wl, fields, tun.waves_per_cu)  # warm-up
    ctx.timing_reset()
    ctx.timing(True)
    steps, sec, pool = B.run_batch(ctx, wl, fields, tun.waves_per_cu, pool=pool)
    ctx.synchronize()
    ctx.timing(False)
    kt = ctx.kernel_time("pool_run")
    kernel_ms = kt[0] / max(kt[1], 1)
    done = pool.request_done.download(np.int32, len(wl.sets))
    complete = bool(all(done[k] == s["size"] for k, s in enumerate(wl.sets)))
    want_pos = pool.position.download(np.float32, 4 * wl.n)
    # the same round captured: init (clearing the request counters) + run, replayed
    t = B.make_tracer(ctx)
    dev = pa.Particles.from_host(ctx, wl.particles)
    rej, res = ctx.alloc(8), ctx.alloc(16)
    with pa.Graph.capture(ctx) as g:
        pa.pool_init_async(ctx, pool, dev, len(fields), rej, zero_requests=True)
        pa.pool_run_async(pool, t, fields, res, n=wl.n, waves_per_cu=tun.waves_per_cu)
    g.launch()  # warm-up replay
    ctx.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    ctx.synchronize()
    gsec = (time.perf_counter() - t0) / reps
    gsteps = int(res.download(np.uint64, 2)[0])
    graph_same = bool(gsteps == steps and pool.position.download(np.float32, 4 * wl.n).tobytes() == want_pos.tobytes())
    g.close()
    # the same batch with every set's curves: bounded passes write each particle's slots into an arena,
    # and each set is finished (prune, index, colour, in the reference's chunks) as soon as it has ended
    ctx.timing_reset()
    ctx.timing(True)
    vt = B.run_batch_vertices(ctx, wl, fields, tun.waves_per_cu)
    ctx.synchronize()
    ctx.timing(False)
    vk = ctx.kernel_time("pool_run")
    pass_ms = vk[0]
    curves = dict(vt, steps_per_s=vt["steps"] / vt["seconds"], pool_run_ms=pass_ms,
                  pool_run_roofline=dict(algorithmic_bytes_per_step=48 + 16,
                                         achieved_gbs=vt["steps"] * 64 / (pass_ms * 1e-3) / 1e9,
                                         frac=vt["steps"] * 64 / (pass_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
                  note="vertices of every set kept and finished on the device set by set (counted prune, "
                       "index, colour per chunk of <= 134,217,727 vertices), then discarded")
    for f in fields:
        f.release()
    # the rate of unrelated-line reads over the same footprint (64-B cell records), what the pool's
    # cell changes cost once they miss the caches
    footprint = sum((g - 1) ** 3 for g in wl.grids) * 64
    scattered = dict(footprint_bytes=footprint, line_bytes=64, read_bytes=48,
                     reads_per_s=scattered_line_rate(pa, ctx, footprint),
                     note="pa_gpu_gather_lines: 48-B reads of pseudo-random 64-B records over the batch's "
                          "cell-record footprint, 8 in flight per lane")
    gather = 48  # BF16: 8 corners x 3 components x 2 B
    return dict(workload="8 curl fields of 48 random modes (64^3..256^3 nodes, BF16), 64 seed sets "
                         "(%d seeds), 100..2000 RK4 steps per set, dt 0.01, one persistent launch" % wl.n,
                particles=wl.n, steps=steps, seconds=sec, steps_per_s=steps / sec, all_sets_complete=complete,
                kernel_ms=kernel_ms,
                roofline=dict(bound="hbm", kernel="k_pool_run<BF16 %s, RK4>" %
                              ("cells" if tun.layout == pa.LAYOUT_CELLS else "nodes"),
                              algorithmic_bytes_per_step=gather, achieved_gbs=steps * gather / (kernel_ms * 1e-3) / 1e9,
                              peak=HBM_PEAK_GBS, frac=steps * gather / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
                graph=dict(seconds_per_round=gsec, steps_per_s=gsteps / gsec, same_results=graph_same),
                scattered_lines=scattered,
                curves=curves,
                tuning=dict(layout="cells" if tun.layout == pa.LAYOUT_CELLS else "nodes",
                            waves_per_cu=tun.waves_per_cu, sample_rates=tun.table))


def bench_serving(pa, ctx, rates=(200.0, 400.0, 800.0, 1600.0), n_requests=60):
    """Serving: one 1024^3 BF16 random-Fourier field (64 modes) resident;
    Poisson arrivals; requests of 256..262144 seeds in boxes of 1-10% of the
    domain; 500 RK4 steps, dt 0.002.  Continuous batching against one
    request at a time on an FP32 copy of the field."""
    from pars_amd import serve as S

    f16 = S.make_field(ctx, pa.BF16).set_layout(pa.LAYOUT_CELLS)
    f32 = S.make_field(ctx, pa.FP32)
    server = S.Server(ctx, [f16], pass_iterations=200)
    resident = S.ResidentServer(ctx, [f16])
    out = dict(workload="1024^3 BF16 curl field of 64 random modes (8.6 GB nodes, served from 64-B cell records) "
                        "resident; Poisson arrivals; 256..262144 seeds per request in 1-10%% boxes; 500 RK4 steps, "
                        "dt 0.002; %d requests per rate; resident = device-resident loop, continuous = host-driven "
                        "passes; baseline: one request at a time on the FP32 node field" % n_requests, rates={})
    warm = S.make_requests(4, 1000.0, seed=1)
    server.serve(warm, keep_endpoints=False)
    resident.serve(S.make_requests(4, 1000.0, seed=1), keep_endpoints=False)
    for rate in rates:
        res_reqs = S.make_requests(n_requests, rate, seed=int(rate))
        rsteps, rwall = resident.serve(res_reqs, keep_endpoints=False)
        reqs = S.make_requests(n_requests, rate, seed=int(rate))
        steps, wall = server.serve(reqs, keep_endpoints=False)
        base = S.make_requests(n_requests, rate, seed=int(rate))
        bsteps, bwall = S.serve_one_at_a_time(ctx, [f32], base, keep_endpoints=False)
        out["rates"]["%g/s" % rate] = dict(
            resident=dict(steps_per_s=rsteps / rwall, requests_per_s=n_requests / rwall,
                          **S.latency_stats(res_reqs)),
            continuous=dict(steps_per_s=steps / wall, requests_per_s=n_requests / wall, **S.latency_stats(reqs)),
            one_at_a_time_fp32=dict(steps_per_s=bsteps / bwall, requests_per_s=n_requests / bwall,
                                    **S.latency_stats(base)))
        log("serving %g/s: %s" % (rate, out["rates"]["%g/s" % rate]))
    # saturated: every request at once
    sat = S.make_requests(n_requests, 1e9, seed=7)
    ssteps, swall = resident.serve(sat, keep_endpoints=False)
    out["saturated_resident"] = dict(steps_per_s=ssteps / swall, requests_per_s=n_requests / swall,
                                     **S.latency_stats(sat))
    resident.close()
    f16.release()
    f32.release()
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--forms", default="fp32-cells-pipelined,fp32-cells,bf16-cells-pipelined,bf16-cells,fp32,bf16,"
                                       "fp16-cells,fp32-cells-consistent,fp32-cells-stages,fp32-cells-lane")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true", help="skip the FTLE, batched and serving workloads")
    ap.add_argument("--workloads", default="ftle,batched,serving", help="which of the other workloads to run")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the fused round's default launch constants")
    ap.add_argument("--launch", default="", help="quad_switch,pair_live,xcd_remap for the fused round (no tuning)")
    ap.add_argument("--calib-bytes", type=int, default=0, help="also run a plain copy of this size (counter calibration)")
    # one GPU's share of the host is 16 CPUs on the boxes this runs on
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the per-rank share on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import pars_amd as pa
    from pars_amd import workloads as W

    ctx = pa.Context(0)
    dev = ctx.device_info()
    log("device:", dev)
    share = W.share(512, 64, 21)
    blocks = [share.local] + list(share.neighbors)
    n_seeds = pa.seed_generator.count(share.local.size, share.local.spacing * np.float32(2))[0]
    assert n_seeds == 262144

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    def storage_of_form(form):
        return {"fp32": pa.FP32, "bf16": pa.BF16, "fp16": pa.FP16}[form.split("-")[0]]

    def layout_of_form(form):
        return pa.LAYOUT_CELLS if "cells" in form.split("-") else pa.LAYOUT_NODES

    storage_of = {f: storage_of_form(f) for f in forms}
    fields = {}
    results = {}
    pipes = {}
    for form in forms:
        st = storage_of[form]
        lay = layout_of_form(form)
        key = (st, lay)
        if key not in fields:
            t0 = time.perf_counter()
            fields[key] = [pa.VectorField.synthesize(ctx, "abc", W.ABC, b.ghosted, b.int_offset, share.h, b.offset,
                                                     b.size, b.spacing, st).set_layout(lay) for b in blocks]
            log("fields %s: %.1f MB in %.2f s" % (form, sum(f.nbytes() for f in fields[key]) / 1e6,
                                                  time.perf_counter() - t0))
        interp = pa.INTERP_CONSISTENT if "consistent" in form.split("-") else pa.INTERP_REFERENCE
        tokens = form.split("-")
        order = "seed" if "seedmajor" in tokens else "chunked" if "quad" in tokens else \
            "step" if "lane" in tokens else "group8"
        engine = pa.ENGINE_LANE if "lane" in tokens else pa.ENGINE_HYBRID if order == "group8" else None
        fused = engine == pa.ENGINE_HYBRID and "stages" not in tokens
        p, res = run_form(pa, ctx, fields[key], share, st, interp, n_seeds, args.steps, args.warmup, order, engine,
                          fused, graph="graph" in tokens, pipelined="pipelined" in tokens, tune=not args.no_tune,
                          launch=[int(x) for x in args.launch.split(",")] if args.launch else None)
        log("%s: %.3f ms/step, %.3e steps/s, trace kernel %s ms, %s" %
            (form, res["ms_per_step"], res["value"], "%.3f" % res["trace_ms"] if res["trace_ms"] else "-",
             res["info"]))
        results[form] = res
        pipes[form] = (p, st, interp)

    if args.calib_bytes:
        a = ctx.alloc(args.calib_bytes)
        b = ctx.alloc(args.calib_bytes)
        for _ in range(3):
            ctx.check(pa.lib().pa_gpu_copy(ctx.handle, C.c_void_p(b.ptr), C.c_void_p(a.ptr),
                                           C.c_size_t(args.calib_bytes)), "copy")
        ctx.synchronize()

    base = results[forms[0]]
    # The headline form's trace kernel as it runs alone: the synchronous twin of a pipelined form (the
    # same kernel on the same workload, with no finish of the previous pass beside it).
    kform = forms[0].replace("-pipelined", "") if forms[0].replace("-pipelined", "") in results else forms[0]
    kres = results[kform]
    steps_per_pass = kres["steps_per_pass"]
    st0 = storage_of[forms[0]]
    alg_bytes = steps_per_pass * (GATHER_BYTES[st0] + STORE_BYTES)
    achieved = alg_bytes / (kres["trace_ms"] * 1e-3) / 1e9
    traffic = load_traffic("trace_union_share_fp32_cells.traffic.json")
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=(traffic or {}).get("bytes_per_launch"),
                    kernel="trace_union, form %s (%s)" % (kform, kernel_name(kform)),
                    algorithmic_bytes_per_step=GATHER_BYTES[st0] + STORE_BYTES,
                    steps_per_launch=steps_per_pass, avg_launch_ms=kres["trace_ms"],
                    effective_gather_gbs=steps_per_pass * GATHER_BYTES[st0] / (kres["trace_ms"] * 1e-3) / 1e9)
    if kform != forms[0] and base["trace_ms"]:
        roofline["in_headline_timing"] = dict(
            form=forms[0], avg_launch_ms=base["trace_ms"],
            frac=alg_bytes / (base["trace_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            note="the same kernel in the pipelined timed region, sharing the chip with the previous pass's finish")
    if traffic and traffic.get("avg_launch_ns"):
        # the committed rocprofv3 summary of the same kernel and workload
        prof_ms = traffic["avg_launch_ns"] * 1e-6
        roofline["profile"] = dict(source=traffic.get("source"), avg_launch_ms=prof_ms,
                                   frac=alg_bytes / (prof_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   duration_ratio=kres["trace_ms"] / prof_ms,
                                   l2_hit_pct=traffic.get("l2_hit_pct"), valu_busy_pct=traffic.get("valu_busy_pct"),
                                   dram_gbs=traffic["bytes_per_launch"] / (prof_ms * 1e-3) / 1e9)

    extra = {}
    for form in forms:
        r = results[form]
        st = storage_of[form]
        tm = r["trace_ms"]
        extra[form] = dict(value=r["value"], ms_per_step=r["ms_per_step"], trace_kernel_ms=tm,
                           launch_tuning=r["tuning"],
                           kernel=kernel_name(form), pass_latency_ms=r["pass_latency_ms"],
                           trace_steps_per_s=r["steps_per_pass"] / (tm * 1e-3) if tm else None,
                           steps_per_pass=r["steps_per_pass"], stage_ms=r["stage_ms"] if tm else None,
                           vs_fp32=r["value"] / base["value"],
                           trace_vs_fp32=(r["steps_per_pass"] / tm) / (base["steps_per_pass"] / base["trace_ms"])
                           if tm else None,
                           effective_gather_gbs=r["steps_per_pass"] * GATHER_BYTES[st] / (tm * 1e-3) / 1e9
                           if tm else None)

    orc = None
    if not args.no_errors or not args.no_cpu_baseline:
        from oracle import oracle as orc

        orc.build()
    if not args.no_errors:
        orc.set_threads(args.cpu_threads)
        for form in forms:
            p, st, interp = pipes[form]
            t0 = time.perf_counter()
            extra[form]["fp64_error"] = endpoint_errors(pa, orc, p, share, st, interp, 32)
            log("errors %s (%.1f s): %s" % (form, time.perf_counter() - t0,
                                            {k: v for k, v in extra[form]["fp64_error"].items() if k[0] != "_"}))
        # the split of a 16-bit form's error: its FP32 twin (same layout, interpolant and engine) against
        # FP64 is the arithmetic floor; the 16-bit form against that twin is the storage part
        for form in forms:
            if storage_of[form] == pa.FP32:
                continue
            twin = "fp32" + form[form.index("-"):] if "-" in form else "fp32"
            if twin in extra and "fp64_error" in extra[twin]:
                extra[form]["storage_error_vs_fp32"] = dict(
                    twin=twin, **storage_error(extra[form]["fp64_error"], extra[twin]["fp64_error"],
                                               float(share.local.size[0])))
                log("storage error %s vs %s: %s" % (form, twin, extra[form]["storage_error_vs_fp32"]))
        for form in forms:
            for k in ("_ends", "_steps"):
                extra[form]["fp64_error"].pop(k, None)
    cpu, parity = None, None
    if not args.no_cpu_baseline:
        # the headline pass once more, kept for the parity check against the oracle
        p0 = pipes[forms[0]][0]
        info = p0.step()
        n0 = info["particles"]
        gpu = dict(form=forms[0], steps=info["steps"], counts=p0.ub["counts"].download(np.int32, n0),
                   vertices=p0.pruned.download(np.float32, 4 * info["vertices"]).reshape(-1, 4),
                   seeds=p0.seeds.download(pa.PARTICLE_DTYPE, n0))
        src = fields[(st0, layout_of_form(forms[0]))]
        ofields = [orc.Field(f.read(), b.offset, b.size, b.spacing) for f, b in zip(src, blocks)]
        t0 = time.perf_counter()
        cpu, parity = cpu_baseline(orc, share, ofields, args.cpu_threads, gpu)
        del gpu, ofields
        log("cpu baseline (%.1f s): %s" % (time.perf_counter() - t0, cpu))
        log("parity of the headline pass against the oracle:", parity)

    # the forms' buffers (two sets for a pipelined form, ~14 GB a set) are not needed past this point
    pipes.clear()
    p0 = None
    import gc
    gc.collect()

    workloads = {}
    if not args.no_workloads:
        chosen = set(w.strip() for w in args.workloads.split(","))
        for name, fn in (("ftle", lambda: bench_ftle(pa, ctx, orc)), ("batched", lambda: bench_batched(pa, ctx)),
                         ("serving", lambda: bench_serving(pa, ctx))):
            if name not in chosen:
                continue
            t0 = time.perf_counter()
            try:
                workloads[name] = fn()
            except Exception as e:  # reported in the line; the headline stands on its own
                workloads[name] = dict(error="%s: %s" % (type(e).__name__, e))
            log("%s (%.1f s): %s" % (name, time.perf_counter() - t0, workloads[name]))

    line = dict(
        metric=METRIC, value=base["value"], unit="particle-steps/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=base["ms_per_step"], higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload="one rank's share: ABC 512^3 split 4x4x4, rank 21 + 6 face neighbours "
                             "(7 x 129^3 ghosted FP32 vec4 blocks, 240 MB of nodes; cell-record layout "
                             "2.1 GB), 262144 seeds stride 2, "
                             "runge_kutta_4 (reference frozen-derivative stepper), dt 0.005, <= 1000 steps, "
                             "union-of-7 continuation, stop |u| < 1e-6; step = seeds + trace + prune + index + "
                             "colour (hsv_constant_s 0.75); passes queued back to back, each finish beside the "
                             "next trace" if "pipelined" in forms[0].split("-") else
                             "one rank's share: ABC 512^3 split 4x4x4, rank 21 + 6 face neighbours "
                             "(7 x 129^3 ghosted FP32 vec4 blocks, 240 MB of nodes; cell-record layout "
                             "2.1 GB), 262144 seeds stride 2, "
                             "runge_kutta_4 (reference frozen-derivative stepper), dt 0.005, <= 1000 steps, "
                             "union-of-7 continuation, stop |u| < 1e-6; step = seeds + trace + prune + index + "
                             "colour (hsv_constant_s 0.75); one pass at a time",
                    seeds=n_seeds, remaining_iterations=R, dt=DT, field="ABC A=sqrt3 B=sqrt2 C=1",
                    form=forms[0], device=dev["name"]),
        pass_latency_ms=dict(kres["pass_latency_ms"], form=kform),
        roofline=roofline, cpu_baseline=cpu, parity=parity, forms=extra, workloads=workloads,
        bf16_vs_fp32=bf16_ratio(forms, extra),
    )
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
