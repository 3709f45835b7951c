"""Autotuner of the fused round's launch constants.

The hybrid engine's speed depends on when a wave moves its last particles
onto quads (quad_switch), on how long FP32 cell records are gathered by lane
pairs (pair_live) and on the XCD-aware workgroup order; its results do not.
tune_round times the caller's own round (pa_gpu_tracer_trace_union_round_async
on the caller's seeds and buffers) for each candidate, checks that every
candidate gives the same step and vertex totals, and leaves the fastest set
on the tracer.  A cache file keeps the choice per device and workload key.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence

import numpy as np

from . import ROUND_RESULT_DTYPE, Context, DeviceBuffer, Particles, particle_tracer

QUAD_SWITCH = (8, 12, 16)
PAIR_LIVE = (0, 16, 32)
XCD_REMAP = (0, 1, 2, 4, 8)


@dataclass
class RoundTuning:
    quad_switch: int
    pair_live: int
    xcd_remap: int
    ms: float
    default_ms: float
    table: Dict[str, float] = field(default_factory=dict)

    def apply(self, tracer: particle_tracer):
        tracer.set_launch(self.quad_switch, self.pair_live)
        tracer.set_xcd_remap(int(self.xcd_remap))

    def as_dict(self):
        return dict(quad_switch=self.quad_switch, pair_live=self.pair_live, xcd_remap=self.xcd_remap, ms=self.ms,
                    default_ms=self.default_ms, table=self.table)


def _label(q, p, x):
    return "q%d/p%d/x%d" % (q, p, x)


def tune_round(ctx: Context, tracer: particle_tracer, seeds: Particles, region_slots: int, buffers: dict,
               result: DeviceBuffer, quad_switch: Sequence[int] = QUAD_SWITCH, pair_live: Sequence[int] = PAIR_LIVE,
               xcd_remap: Sequence[int] = XCD_REMAP, reps: int = 3, cache: Optional[str] = None,
               key: Optional[str] = None) -> RoundTuning:
    """Time every (quad_switch, pair_live, xcd_remap) on the caller's round; the median of `reps` timed
    rounds after one warm-up each.  buffers as Tracer.round_buffers (out_vertices, out_indices,
    out_colors included)."""
    if cache and key and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        if key in known:
            k = known[key]
            tun = RoundTuning(k["quad_switch"], k["pair_live"], k["xcd_remap"], k["ms"], k["default_ms"], k["table"])
            tun.apply(tracer)
            return tun
    table = {}
    totals = None
    best = None
    host = np.zeros(1, ROUND_RESULT_DTYPE)
    for q, p, x in itertools.product(quad_switch, pair_live, xcd_remap):
        tracer.set_launch(q, p)
        tracer.set_xcd_remap(int(x))
        times = []
        for r in range(reps + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            tracer.trace_union_round_async(seeds, region_slots, buffers, result, 2, 0.75)
            ctx.synchronize()
            if r:
                times.append(time.perf_counter() - t0)
        host[:] = result.download(ROUND_RESULT_DTYPE, 1)
        tot = (int(host["total_steps"][0]), int(host["n_vertices"][0]), int(host["n_indices"][0]))
        if totals is None:
            totals = tot
        elif tot != totals:
            raise RuntimeError("launch constants %s changed the round's totals %s -> %s" % (_label(q, p, x), totals,
                                                                                          tot))
        ms = 1e3 * float(np.median(times))
        table[_label(q, p, x)] = ms
        if best is None or ms < best[3]:
            best = (q, p, x, ms)
    default = table.get(_label(16, 16, 1), best[3])  # the tracer's defaults
    tun = RoundTuning(best[0], best[1], best[2], best[3], default, table)
    tun.apply(tracer)
    if cache and key:
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[key] = tun.as_dict()
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return tun
