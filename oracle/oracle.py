"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes front end of ``liboracle.so``, the CPU restatement of pars'
per-process advection path (see ``pa_oracle.h`` for the reference
citations and the pinning status of every function).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PARTICLE_DTYPE = np.dtype(
    [
        ("position", "<f4", (4,)),
        ("remaining_iterations", "<i4"),
        ("vector_field_index", "<i4"),
        ("original_rank", "<i4"),
        ("original_voxel", "<i4", (3,)),
        ("reserved_", "<i4", (2,)),
    ]
)
assert PARTICLE_DTYPE.itemsize == 48

HANDOFF_DTYPE = np.dtype(
    [("particle", PARTICLE_DTYPE), ("kind", "<i4"), ("slot", "<i4"), ("source", "<i8")]
)
assert HANDOFF_DTYPE.itemsize == 64

EULER, RK4, CLASSIC_RK4 = 0, 2, 100
RK4_DT_OVER_6 = 102  # runge_kutta4_classic's coefficient rounding (risk measurement only)
STOP_ISZERO, STOP_NORM = 0, 1
INTERP_REFERENCE, INTERP_CONSISTENT = 0, 1
TERMINATION = np.array([-2.0, -2.0, -2.0, 0.0], np.float32)
INVALID = np.array([-1.0, -1.0, -1.0, 0.0], np.float32)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


class _Field(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("dims", C.c_int32 * 3),
        ("offset", C.c_float * 3),
        ("size", C.c_float * 3),
        ("spacing", C.c_float * 3),
    ]


class _Field64(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("dims", C.c_int32 * 3),
        ("offset", C.c_double * 3),
        ("size", C.c_double * 3),
        ("spacing", C.c_double * 3),
    ]


class _Config(C.Structure):
    _fields_ = [
        ("local", C.POINTER(_Field)),
        ("neighbors", C.POINTER(_Field) * 6),
        ("neighbor_rank", C.c_int32 * 6),
        ("neighbor_face_mask", C.c_int32 * 6),
        ("step_size", C.c_float),
        ("integrator", C.c_int32),
        ("stop_rule", C.c_int32),
        ("stop_eps", C.c_float),
        ("interp", C.c_int32),
    ]


class _Config64(C.Structure):
    _fields_ = [
        ("local", C.POINTER(_Field64)),
        ("neighbors", C.POINTER(_Field64) * 6),
        ("neighbor_rank", C.c_int32 * 6),
        ("neighbor_face_mask", C.c_int32 * 6),
        ("step_size", C.c_double),
        ("integrator", C.c_int32),
        ("stop_rule", C.c_int32),
        ("stop_eps", C.c_double),
        ("interp", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        _lib.orc_seed_count.restype = C.c_int64
        _lib.orc_prune.restype = C.c_size_t
        _lib.orc_index_generate.restype = C.c_int64
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f3(vals, ctype=C.c_float):
    return (ctype * 3)(*[float(v) for v in vals])


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def hardware_concurrency() -> int:
    return int(lib().orc_hardware_concurrency())


class Field:
    """A block as the reference's vector_field: xyz data [d0][d1][d2][3]."""

    def __init__(self, data, offset, size, spacing):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        assert self.data.ndim == 4 and self.data.shape[3] == 3
        self.offset = np.asarray(offset, np.float32)
        self.size = np.asarray(size, np.float32)
        self.spacing = np.asarray(spacing, np.float32)
        self._s = _Field(
            _p(self.data).value,
            (C.c_int32 * 3)(*self.data.shape[:3]),
            _f3(self.offset),
            _f3(self.size),
            _f3(self.spacing),
        )

    @property
    def ref(self):
        return C.byref(self._s)


class Field64:
    def __init__(self, data, offset, size, spacing):
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self.offset = np.asarray(offset, np.float64)
        self.size = np.asarray(size, np.float64)
        self.spacing = np.asarray(spacing, np.float64)
        self._s = _Field64(
            _p(self.data).value,
            (C.c_int32 * 3)(*self.data.shape[:3]),
            _f3(self.offset, C.c_double),
            _f3(self.size, C.c_double),
            _f3(self.spacing, C.c_double),
        )

    @property
    def ref(self):
        return C.byref(self._s)


def _config(local, neighbors=None, neighbor_rank=None, face_masks=None, dt=0.01, integrator=RK4,
            stop_rule=STOP_ISZERO, stop_eps=0.0, f64=False, interp=INTERP_REFERENCE):
    cfg = (_Config64 if f64 else _Config)()
    ftype = _Field64 if f64 else _Field
    cfg.local = C.pointer(local._s)
    neighbors = neighbors or [None] * 6
    for k in range(6):
        if neighbors[k] is not None:
            cfg.neighbors[k] = C.pointer(neighbors[k]._s)
        else:
            cfg.neighbors[k] = C.POINTER(ftype)()
    ranks = neighbor_rank if neighbor_rank is not None else [-1] * 6
    masks = face_masks if face_masks is not None else [0] * 6
    for k in range(6):
        cfg.neighbor_rank[k] = int(ranks[k])
        cfg.neighbor_face_mask[k] = int(masks[k])
    cfg.step_size = dt
    cfg.integrator = integrator
    cfg.stop_rule = stop_rule
    cfg.stop_eps = stop_eps
    cfg.interp = interp
    cfg._keep = (local, neighbors)
    return cfg


def seed_count(size, stride):
    dims = (C.c_int32 * 3)()
    n = lib().orc_seed_count(_f3(size), _f3(stride), dims)
    return int(n), tuple(dims)


def seed_generate(offset, size, stride, remaining_iterations, rank=0):
    n, _ = seed_count(size, stride)
    out = np.zeros(max(n, 0), PARTICLE_DTYPE)
    if n > 0:
        lib().orc_seed_generate(_f3(offset), _f3(size), _f3(stride), C.c_int32(remaining_iterations),
                                C.c_int32(rank), _p(out))
    return out


def contains(field, p):
    p = np.ascontiguousarray(p, np.float32)
    return bool(lib().orc_contains(field.ref, _p(p)))


def interpolate(field, p, interp=INTERP_REFERENCE):
    p = np.ascontiguousarray(p, np.float32)
    out = np.zeros(3, np.float32)
    fn = lib().orc_interpolate if interp == INTERP_REFERENCE else lib().orc_interpolate_consistent
    fn(field.ref, _p(p), _p(out))
    return out


def step(integrator, p, v, dt):
    p = np.ascontiguousarray(p, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(4, np.float32)
    lib().orc_step(C.c_int32(integrator), _p(p), _p(v), C.c_float(dt), _p(out))
    return out


def round_info(particles, curve_offset=0):
    out = (C.c_uint64 * 5)()
    rc = lib().orc_round_info(_p(particles), C.c_size_t(len(particles)), C.c_uint64(curve_offset), out)
    if rc != 0:
        raise ValueError("round_info of an empty or non-positive particle set")
    return tuple(int(x) for x in out)


def chunk_sizes(info):
    _, vpic, vpic_last, _, count = info
    return [vpic if c != count - 1 else vpic_last for c in range(count)]


def trace(particles, local, neighbors=None, neighbor_rank=None, dt=0.01, integrator=RK4,
          stop_rule=STOP_ISZERO, stop_eps=0.0, curve_offset=0, interp=INTERP_REFERENCE, fill=True):
    """Reference round: compute_round_info + allocate + initialize + trace.

    Returns (info, chunks, vertex_counts, handoffs, total_steps)."""
    particles = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    info = round_info(particles, curve_offset)
    chunks = [(np.tile(INVALID, (sz, 1)) if fill else np.empty((sz, 4), np.float32)) for sz in chunk_sizes(info)]
    ptrs = (C.c_void_p * max(1, len(chunks)))(*[_p(c).value for c in chunks])
    counts = np.zeros(len(particles), np.int32)
    hand = np.zeros(len(particles), HANDOFF_DTYPE)
    nh = C.c_size_t()
    steps = C.c_uint64()
    cfg = _config(local, neighbors, neighbor_rank, None, dt, integrator, stop_rule, stop_eps, interp=interp)
    rc = lib().orc_trace(C.byref(cfg), _p(particles), C.c_size_t(len(particles)),
                         (C.c_uint64 * 5)(*info), ptrs, _p(counts), _p(hand), C.byref(nh), C.byref(steps))
    if rc != 0:
        raise ValueError("orc_trace rejected its input (%d)" % rc)
    return info, chunks, counts, hand[: nh.value], int(steps.value)


def trace_union(particles, local, neighbors, neighbor_rank, face_masks, dt=0.01, integrator=RK4,
                stop_rule=STOP_ISZERO, stop_eps=0.0, vertex_capacity=None, segment_capacity=None,
                interp=INTERP_REFERENCE):
    """Union-of-7 emulation: returns dict with vertices, segment_seed,
    segment_block, segment_begin, segment_length, total_steps."""
    particles = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    n = len(particles)
    rmax = int(particles["remaining_iterations"].max()) if n else 0
    vcap = vertex_capacity or max(1, n * (rmax + 16))
    scap = segment_capacity or max(1, n * 16)
    verts = np.zeros((vcap, 4), np.float32)
    seed = np.zeros(scap, np.int64)
    block = np.zeros(scap, np.int32)
    begin = np.zeros(scap, np.int64)
    length = np.zeros(scap, np.int32)
    nv = C.c_size_t()
    ns = C.c_size_t()
    steps = C.c_uint64()
    cfg = _config(local, neighbors, neighbor_rank, face_masks, dt, integrator, stop_rule, stop_eps, interp=interp)
    rc = lib().orc_trace_union(C.byref(cfg), _p(particles), C.c_size_t(n), _p(verts), C.c_size_t(vcap),
                               C.byref(nv), _p(seed), _p(block), _p(begin), _p(length), C.c_size_t(scap),
                               C.byref(ns), C.byref(steps))
    if rc != 0:
        raise ValueError("orc_trace_union failed (%d)" % rc)
    m = ns.value
    return dict(vertices=verts[: nv.value], segment_seed=seed[:m], segment_block=block[:m],
                segment_begin=begin[:m], segment_length=length[:m], total_steps=int(steps.value))


def endpoints(particles, local, neighbors=None, neighbor_rank=None, face_masks=None, dt=0.01, integrator=RK4,
              union=False, stop_rule=STOP_ISZERO, stop_eps=0.0, interp=INTERP_REFERENCE):
    particles = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    out = np.zeros((len(particles), 4), np.float32)
    steps = np.zeros(len(particles), np.int32)
    cfg = _config(local, neighbors, neighbor_rank, face_masks, dt, integrator, stop_rule, stop_eps, interp=interp)
    lib().orc_endpoints(C.byref(cfg), _p(particles), C.c_size_t(len(particles)), C.c_int(int(union)), _p(out),
                        _p(steps))
    return out, steps


def endpoints_f64(positions, remaining_iterations, local, neighbors=None, neighbor_rank=None, face_masks=None,
                  dt=0.01, integrator=RK4, union=False, stop_rule=STOP_ISZERO, stop_eps=0.0):
    pos = np.ascontiguousarray(positions, np.float64)
    if pos.ndim != 2 or pos.shape[1] != 4:
        raise ValueError("endpoints_f64: positions are (n, 4) x, y, z, w")
    out = np.zeros((len(pos), 4), np.float64)
    steps = np.zeros(len(pos), np.int32)
    cfg = _config(local, neighbors, neighbor_rank, face_masks, dt, integrator, stop_rule, stop_eps, f64=True)
    lib().orc_endpoints_f64(C.byref(cfg), _p(pos), C.c_size_t(len(pos)), C.c_int32(remaining_iterations),
                            C.c_int(int(union)), _p(out), _p(steps))
    return out, steps


def prune(vertices):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
    out = np.zeros_like(v)
    m = lib().orc_prune(_p(v), C.c_size_t(len(v)), _p(out))
    return out[:m]


def index_generate(vertices):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
    out = np.zeros(max(1, len(v)), np.int32)
    m = lib().orc_index_generate(_p(v), C.c_size_t(len(v)), _p(out))
    if m < 0:
        raise ValueError("index generation of an empty vertex array")
    return out[:m]


def color_generate(vertices, indices, mode=4, parameter=0.5):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
    idx = np.ascontiguousarray(indices, np.int32)
    colors = np.zeros_like(v)
    lib().orc_color_generate(_p(v), C.c_size_t(len(v)), _p(idx), C.c_size_t(len(idx)), C.c_int32(mode),
                             C.c_float(parameter), _p(colors))
    return colors


def advect(particles, local, neighbor_rank=None, dt=0.01, integrator=RK4, stop_rule=STOP_ISZERO, stop_eps=0.0,
           interp=INTERP_REFERENCE):
    p = np.array(particles, PARTICLE_DTYPE, copy=True)
    state = np.zeros(len(p), np.int32)
    slot = np.zeros(len(p), np.int32)
    steps = C.c_uint64()
    cfg = _config(local, None, neighbor_rank, None, dt, integrator, stop_rule, stop_eps, interp=interp)
    lib().orc_advect(C.byref(cfg), _p(p), C.c_size_t(len(p)), _p(state), _p(slot), C.byref(steps))
    return p, state, slot, int(steps.value)


def flow_map(field, block_size, iterations, resolution_scale=1.0, rank=0, integrator=RK4, dt=0.01):
    dims = [int(np.float32(b) * np.float32(resolution_scale)) for b in block_size]
    out = np.zeros((dims[0], dims[1], dims[2], 3), np.float32)
    dims_out = (C.c_int32 * 3)()
    steps = C.c_uint64()
    lib().orc_flow_map(field.ref, (C.c_int32 * 3)(*block_size), C.c_uint64(iterations),
                       C.c_float(resolution_scale), C.c_int32(rank), C.c_int32(integrator), C.c_float(dt),
                       _p(out), dims_out, C.byref(steps))
    assert tuple(dims_out) == tuple(dims)
    return out, int(steps.value)


def ftle(flow_map_xyz, spacing, time, mode=0):
    fm = np.ascontiguousarray(flow_map_xyz, np.float32)
    dims = fm.shape[:3]
    out = np.zeros(dims, np.float32)
    lmax = np.zeros(dims, np.float64)
    lib().orc_ftle(_p(fm), (C.c_int32 * 3)(*dims), _f3(spacing), C.c_double(time), C.c_int32(mode), _p(out),
                   _p(lmax))
    return out, lmax


def eigh3(a):
    a = np.asarray(a, np.float64)
    sym = np.ascontiguousarray([a[0, 0], a[0, 1], a[0, 2], a[1, 1], a[1, 2], a[2, 2]], np.float64)
    ev = np.zeros(3, np.float64)
    lib().orc_eigh3(_p(sym), _p(ev))
    return ev


def abc_field_f64(dims, origin, h, A=np.sqrt(3.0), B=np.sqrt(2.0), Cc=1.0):
    out = np.zeros(tuple(dims) + (3,), np.float64)
    lib().orc_abc_field_f64((C.c_int32 * 3)(*dims), (C.c_int32 * 3)(*origin), C.c_double(h), C.c_double(A),
                            C.c_double(B), C.c_double(Cc), _p(out))
    return out


def fourier_field_f64(dims, origin, h, modes):
    modes = np.ascontiguousarray(modes, np.float64)
    out = np.zeros(tuple(dims) + (3,), np.float64)
    lib().orc_fourier_field_f64((C.c_int32 * 3)(*dims), (C.c_int32 * 3)(*origin), C.c_double(h), _p(modes),
                                C.c_int32(len(modes)), _p(out))
    return out
