"""pars_amd — MI355X (gfx950) implementation of pars' per-process advection path.

Python mirror of the reference's ``pa`` stage interface (acdemiralp/pars,
``pa/include/pa/stages/*.hpp``) over the C ABI of ``libpa_gpu.so``
(``include/pa_gpu.h``).  Stage and method names follow the reference:
``seed_generator.generate``, ``particle_tracer`` (``set_*``,
``compute_round_info``, ``allocate``, ``initialize``/``trace``, ``prune``),
``index_generator.generate``, ``color_generator.generate``,
``flow_map_generator`` and ``ftle_map_generator``.

All data stays in device memory (``DeviceBuffer``); the HIP library is the
only compute path.  Importing this package fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpa_gpu.so")

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

# pa_status
OK, E_INVALID, E_UNSUPPORTED, E_DEVICE, E_NOMEM, E_CAPACITY, E_STATE = range(7)
# pa_storage
FP32, BF16, FP16 = 0, 1, 2
# pa_integrator (pa::variant_integrator order) + extension
EULER, MODIFIED_MIDPOINT, RUNGE_KUTTA_4, CASH_KARP_54, DORMAND_PRINCE_5, FEHLBERG_78, AB2, ABM2 = range(8)
CLASSIC_RK4 = 100
INTEGRATORS = {
    "euler": EULER, "modified_midpoint": MODIFIED_MIDPOINT, "runge_kutta_4": RUNGE_KUTTA_4,
    "runge_kutta_cash_karp_54": CASH_KARP_54, "runge_kutta_dormand_prince_5": DORMAND_PRINCE_5,
    "runge_kutta_fehlberg_78": FEHLBERG_78, "adams_bashforth_2": AB2, "adams_bashforth_moulton_2": ABM2,
    "classic_rk4": CLASSIC_RK4,
}
STOP_ISZERO, STOP_NORM = 0, 1
COLOR_HSL_CONSTANT_S, COLOR_HSL_CONSTANT_L, COLOR_HSV_CONSTANT_S, COLOR_HSV_CONSTANT_V, COLOR_RGB = range(5)
COLOR_MAGNITUDE = 100
INTERP_REFERENCE, INTERP_CONSISTENT = 0, 1
LAYOUT_NODES, LAYOUT_CELLS = 0, 1
ENGINE_LANE, ENGINE_COOPERATIVE, ENGINE_QUAD, ENGINE_HYBRID = 0, 1, 2, 3
SLOTS_SEED_MAJOR, SLOTS_STEP_MAJOR, SLOTS_CHUNKED, SLOTS_GROUP8 = 0, 1, 2, 3
SLOT_ORDERS = {"seed": SLOTS_SEED_MAJOR, "step": SLOTS_STEP_MAJOR, "chunked": SLOTS_CHUNKED,
               "group8": SLOTS_GROUP8}
TERMINATION_VERTEX = np.array([-2.0, -2.0, -2.0, 0.0], np.float32)
INVALID_VERTEX = np.array([-1.0, -1.0, -1.0, 0.0], np.float32)
INVALID_INDEX = -1


class PaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("pa_gpu status %d: %s" % (status, message))
        self.status = status


class RoundInfo(C.Structure):
    """particle_tracer::round_info (particle_tracer.hpp:26-37) without the maps."""

    _fields_ = [
        ("maximum_remaining_iterations", C.c_uint64),
        ("vertices_per_integral_curve", C.c_uint64),
        ("vertices_per_integral_curve_last", C.c_uint64),
        ("integral_curve_offset", C.c_uint64),
        ("integral_curve_count", C.c_uint64),
    ]

    def chunk_sizes(self):
        n = int(self.integral_curve_count)
        return [int(self.vertices_per_integral_curve) if c != n - 1 else int(self.vertices_per_integral_curve_last)
                for c in range(n)]


class _Handoff(C.Structure):
    _fields_ = [
        ("d_particles", C.c_void_p),
        ("d_sources", C.c_void_p),
        ("capacity", C.c_size_t),
        ("counts", (C.c_size_t * 6) * 2),
    ]


_lib = None


def lib():
    """The loaded libpa_gpu.so.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("pars_amd: %s is missing; run `python -m pars_amd.build` (or __graft_entry__.build())"
                              % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.pa_gpu_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


class Context:
    """A device context (pa_gpu_ctx): one HIP stream on one GPU."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        s = lib().pa_gpu_ctx_create(C.c_int(device), C.byref(self._h))
        if s != OK:
            raise PaError(s, "pa_gpu_ctx_create(%d) failed (no usable GPU?)" % device)

    @property
    def handle(self):
        return self._h

    def check(self, status: int, what: str = ""):
        if status != OK:
            msg = lib().pa_gpu_last_error(self._h).decode()
            raise PaError(status, (what + ": " if what else "") + msg)

    def synchronize(self):
        self.check(lib().pa_gpu_synchronize(self._h), "synchronize")

    def round_join(self):
        """Order the context's stream after every pipelined round's finish (pa_gpu_round_join)."""
        self.check(lib().pa_gpu_round_join(self._h), "round_join")

    def close(self):
        if self._h:
            lib().pa_gpu_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        free = C.c_size_t()
        total = C.c_size_t()
        self.check(lib().pa_gpu_device_info(self._h, name, C.c_size_t(256), C.byref(cu), C.byref(free),
                                            C.byref(total)))
        return dict(name=name.value.decode(), compute_units=cu.value, free_bytes=free.value, total_bytes=total.value)

    # timing -------------------------------------------------------------
    def timing(self, enable: bool = True):
        self.check(lib().pa_gpu_timing_enable(self._h, C.c_int(int(enable))))

    def timing_reset(self):
        self.check(lib().pa_gpu_timing_reset(self._h))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_uint64()
        self.check(lib().pa_gpu_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # memory ---------------------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def upload(self, array: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(array)
        b = DeviceBuffer(self, a.nbytes)
        b.upload(a)
        return b


class DeviceBuffer:
    """Device allocation owned by Python (pa_gpu_malloc / pa_gpu_free)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self._p = C.c_void_p()
        ctx.check(lib().pa_gpu_malloc(ctx.handle, C.c_size_t(max(self.nbytes, 16)), C.byref(self._p)), "malloc")

    @property
    def ptr(self) -> int:
        return self._p.value

    def at(self, byte_offset: int) -> int:
        return self._p.value + int(byte_offset)

    def upload(self, array: np.ndarray, byte_offset: int = 0):
        a = np.ascontiguousarray(array)
        assert byte_offset + a.nbytes <= self.nbytes
        self.ctx.check(lib().pa_gpu_memcpy_h2d(self.ctx.handle, C.c_void_p(self.at(byte_offset)),
                                               a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)), "h2d")

    def download(self, dtype, count: Optional[int] = None, byte_offset: int = 0, shape=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - byte_offset) // dtype.itemsize
        out = np.empty(count, dtype)
        if count:
            self.ctx.check(lib().pa_gpu_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                                   C.c_void_p(self.at(byte_offset)),
                                                   C.c_size_t(count * dtype.itemsize)), "d2h")
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self._p.value:
            lib().pa_gpu_free(self.ctx.handle, self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class VectorField:
    """A block resident in HBM (pa::vector_field, vector_field.hpp:14-24)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, dims, offset, size, spacing, storage):
        self.ctx = ctx
        self._h = handle
        self.dims = tuple(int(d) for d in dims)
        self.offset = np.asarray(offset, np.float32)
        self.size = np.asarray(size, np.float32)
        self.spacing = np.asarray(spacing, np.float32)
        self.storage = storage

    @classmethod
    def from_host(cls, ctx: Context, xyz: np.ndarray, offset, size, spacing, storage=FP32):
        """Upload [d0][d1][d2][3] FP32 data (data_io.cpp:91-115 replacement)."""
        a = np.ascontiguousarray(xyz, np.float32)
        assert a.ndim == 4 and a.shape[3] == 3
        h = C.c_void_p()
        ctx.check(lib().pa_gpu_field_create(ctx.handle, a.ctypes.data_as(C.c_void_p), _i3(a.shape[:3]), _f3(offset),
                                            _f3(size), _f3(spacing), C.c_int(storage), C.byref(h)), "field_create")
        return cls(ctx, h, a.shape[:3], offset, size, spacing, storage)

    @classmethod
    def load(cls, ctx: Context, path: str, offset, ghosted, block, storage=FP32):
        """data_io::load_vector_field (data_io.cpp:91-115): one rank's ghosted HDF5 hyperslab straight
        into device bricks; offset is the rank's integer offset, ghosted and block its node counts."""
        _, off, size, sp = read_vector_block(path, offset, ghosted, block, want_data=False)
        h = C.c_void_p()
        ctx.check(lib().pa_gpu_field_load(ctx.handle, path.encode(), (C.c_int64 * 3)(*[int(x) for x in offset]),
                                          _i3(ghosted), _i3(block), C.c_int(storage), C.byref(h)), "field_load")
        return cls(ctx, h, ghosted, off, size, sp, storage)

    @classmethod
    def synthesize(cls, ctx: Context, kind: str, params, dims, origin, h, offset, size, spacing, storage=FP32):
        """ABC flow (params A, B, C) or Fourier modes (rows k0 k1 k2 c0 c1 c2 phi) on the device."""
        p = np.ascontiguousarray(params, np.float64).ravel()
        k = {"abc": 0, "fourier": 1}[kind]
        hh = C.c_void_p()
        ctx.check(lib().pa_gpu_field_synthesize(ctx.handle, C.c_int(k), p.ctypes.data_as(C.c_void_p),
                                                C.c_int32(len(p)), _i3(dims), _i3(origin), C.c_double(h),
                                                _f3(offset), _f3(size), _f3(spacing), C.c_int(storage),
                                                C.byref(hh)), "field_synthesize")
        return cls(ctx, hh, dims, offset, size, spacing, storage)

    @property
    def handle(self):
        return self._h

    def read(self) -> np.ndarray:
        out = np.empty(self.dims + (3,), np.float32)
        self.ctx.check(lib().pa_gpu_field_read(self.ctx.handle, self._h, out.ctypes.data_as(C.c_void_p)),
                       "field_read")
        return out

    def set_layout(self, layout: int):
        """LAYOUT_CELLS: one record of the 8 corners per cell (one cache line per gather)."""
        self.ctx.check(lib().pa_gpu_field_set_layout(self.ctx.handle, self._h, C.c_int(layout)), "set_layout")
        self.layout = layout
        return self

    def interpolate(self, points: np.ndarray, interp: int = INTERP_REFERENCE) -> np.ndarray:
        """vector_field::contains + interpolate at points (n x 4) -> (n x 4): u, v, w, inside."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
        n = len(pts)
        d_in = self.ctx.upload(pts)
        d_out = self.ctx.alloc(max(n, 1) * 16)
        self.ctx.check(lib().pa_gpu_field_interpolate(self.ctx.handle, self._h, C.c_int(interp),
                                                      C.c_void_p(d_in.ptr), C.c_size_t(n), C.c_void_p(d_out.ptr)),
                       "field_interpolate")
        return d_out.download(np.float32, 4 * n).reshape(n, 4)

    def nbytes(self) -> int:
        b = C.c_size_t()
        lib().pa_gpu_field_bytes(self._h, C.byref(b))
        return b.value

    def release(self):
        if self._h and self._h.value:
            lib().pa_gpu_field_destroy(self.ctx.handle, self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Particles:
    """A device array of pa_particle (48 B each)."""

    def __init__(self, buf: DeviceBuffer, n: int):
        self.buf = buf
        self.n = int(n)

    @property
    def ptr(self):
        return self.buf.ptr

    @classmethod
    def from_host(cls, ctx: Context, particles: np.ndarray):
        p = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        return cls(ctx.upload(p) if len(p) else ctx.alloc(48), len(p))

    def download(self) -> np.ndarray:
        return self.buf.download(PARTICLE_DTYPE, self.n)


class seed_generator:
    """pa::seed_generator (seed_generator.hpp:10-14; seed_generator.cpp:9-21)."""

    @staticmethod
    def count(size, stride):
        dims = (C.c_int32 * 3)()
        n = C.c_size_t()
        s = lib().pa_gpu_seed_count(_f3(size), _f3(stride), dims, C.byref(n))
        if s != OK:
            raise PaError(s, "seed_count")
        return int(n.value), tuple(dims)

    @staticmethod
    def generate(ctx: Context, offset, size, stride, remaining_iterations: int, rank: int = 0) -> Particles:
        n, _ = seed_generator.count(size, stride)
        buf = ctx.alloc(max(n, 1) * 48)
        got = C.c_size_t()
        ctx.check(lib().pa_gpu_seed_generate(ctx.handle, _f3(offset), _f3(size), _f3(stride),
                                             C.c_int32(remaining_iterations), C.c_int32(rank), C.c_void_p(buf.ptr),
                                             C.c_size_t(max(n, 1)), C.byref(got)), "seed_generate")
        return Particles(buf, got.value)


class particle_tracer:
    """pa::particle_tracer (particle_tracer.hpp:39-62) on the GPU."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx.check(lib().pa_gpu_tracer_create(ctx.handle, C.byref(self._h)), "tracer_create")
        self._keep = []

    def __del__(self):
        try:
            lib().pa_gpu_tracer_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_local_vector_field(self, field: VectorField):
        self._local = field
        self.ctx.check(lib().pa_gpu_tracer_set_local_vector_field(self._h, field.handle if field else None))

    def set_neighbor_vector_fields(self, fields: Sequence[Optional[VectorField]], neighbor_rank: Sequence[int]):
        arr = (C.c_void_p * 6)(*[(f.handle.value if f is not None else None) for f in fields])
        self._nbrs = list(fields)
        self.ctx.check(lib().pa_gpu_tracer_set_neighbor_vector_fields(self._h, arr, (C.c_int32 * 6)(*neighbor_rank)))

    def set_neighbor_face_masks(self, masks: Sequence[int]):
        self.ctx.check(lib().pa_gpu_tracer_set_neighbor_face_masks(self._h, (C.c_int32 * 6)(*masks)))

    def set_integrator(self, integrator):
        if isinstance(integrator, str):
            integrator = INTEGRATORS[integrator]
        self.ctx.check(lib().pa_gpu_tracer_set_integrator(self._h, C.c_int(integrator)), "set_integrator")

    def set_step_size(self, step_size: float):
        self.ctx.check(lib().pa_gpu_tracer_set_step_size(self._h, C.c_float(step_size)))

    def set_engine(self, engine: int, waves: int = 0):
        """ENGINE_LANE (default), ENGINE_COOPERATIVE, ENGINE_QUAD (union trace, chunked slots) or
        ENGINE_HYBRID (union trace, step-major slots); waves = resident waves of the cooperative
        engine (0 = automatic)."""
        self.ctx.check(lib().pa_gpu_tracer_set_engine(self._h, C.c_int(engine), C.c_int32(waves)), "set_engine")

    def set_xcd_remap(self, order):
        """XCD workgroup order: 0 hardware order, 1 contiguous eighths, 2..64 runs of that many workgroups."""
        self.ctx.check(lib().pa_gpu_tracer_set_xcd_remap(self._h, C.c_int32(int(order))), "set_xcd_remap")

    def set_launch(self, quad_switch: int = 16, pair_live: int = 16):
        """Hybrid engine launch constants (pa_gpu_tracer_set_launch); results do not depend on them."""
        self.ctx.check(lib().pa_gpu_tracer_set_launch(self._h, C.c_int32(quad_switch), C.c_int32(pair_live)),
                       "set_launch")

    def set_interpolation(self, interp: int):
        self.ctx.check(lib().pa_gpu_tracer_set_interpolation(self._h, C.c_int(interp)), "set_interpolation")

    def set_stop_rule(self, rule: int, eps: float = 0.0):
        self.ctx.check(lib().pa_gpu_tracer_set_stop_rule(self._h, C.c_int(rule), C.c_float(eps)))

    # sub-methods (particle_tracer.hpp:55-61) ---------------------------------
    def compute_round_info(self, particles: Particles, integral_curve_offset: int = 0) -> RoundInfo:
        info = RoundInfo()
        self.ctx.check(lib().pa_gpu_tracer_compute_round_info(self._h, C.c_void_p(particles.ptr),
                                                              C.c_size_t(particles.n),
                                                              C.c_size_t(integral_curve_offset), C.byref(info)),
                       "compute_round_info")
        return info

    def allocate(self, info: RoundInfo, fill_invalid: bool = True):
        chunks = []
        for sz in info.chunk_sizes():
            b = self.ctx.alloc(max(sz, 1) * 16)
            if fill_invalid and sz:
                self.ctx.check(lib().pa_gpu_fill_invalid(self.ctx.handle, C.c_void_p(b.ptr), C.c_size_t(sz)))
            chunks.append((b, sz))
        return chunks

    def trace(self, particles: Particles, chunks, info: RoundInfo, want_counts=True, want_handoff=True):
        """initialize + trace of one round.  Returns dict(counts, handoff, handoff_counts, total_steps)."""
        table = (C.c_void_p * max(1, len(chunks)))(*[c[0].ptr for c in chunks])
        counts = self.ctx.alloc(max(particles.n, 1) * 4) if want_counts else None
        hand = None
        h = None
        if want_handoff:
            hand = self.ctx.alloc(max(particles.n, 1) * 48)
            src = self.ctx.alloc(max(particles.n, 1) * 8)
            h = _Handoff(hand.ptr, src.ptr, max(particles.n, 1))
        steps = C.c_uint64()
        self.ctx.check(lib().pa_gpu_tracer_trace(self._h, C.c_void_p(particles.ptr), C.c_size_t(particles.n),
                                                 C.byref(info), table,
                                                 C.c_void_p(counts.ptr) if counts else None,
                                                 C.byref(h) if h is not None else None, C.byref(steps)), "trace")
        out = dict(total_steps=int(steps.value), counts=counts)
        if h is not None:
            hc = np.array([[h.counts[k][s] for s in range(6)] for k in range(2)], np.int64)
            tot = int(hc.sum())
            out["handoff_counts"] = hc
            out["handoff"] = hand.download(PARTICLE_DTYPE, tot) if tot else np.zeros(0, PARTICLE_DTYPE)
            out["handoff_sources"] = src.download(np.int64, tot) if tot else np.zeros(0, np.int64)
        return out

    def trace_union(self, particles: Particles, region_slots: int, buffers=None, endpoints=False, order="seed"):
        """Union-of-7 trace; buffers (from union_buffers) may be reused across calls.

        order "seed": slot s of seed p at p * region_slots + s; "step": at s * n + p;
        "chunked" (quad engine): 8-slot component-planar chunks per seed; "group8" (hybrid
        engine): at ((p // 8) * region_slots + s) * 8 + p % 8."""
        b = buffers or self.union_buffers(particles.n, region_slots, endpoints)
        ncont = C.c_size_t()
        steps = C.c_uint64()
        self.ctx.check(lib().pa_gpu_tracer_trace_union_layout(
            self._h, C.c_void_p(particles.ptr), C.c_size_t(particles.n), C.c_size_t(region_slots),
            C.c_int(SLOT_ORDERS[order]), C.c_void_p(b["vertices"].ptr), C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.byref(ncont),
            C.c_void_p(b["endpoints"].ptr) if b.get("endpoints") else None,
            C.c_void_p(b["seed_steps"].ptr) if b.get("seed_steps") else None, C.byref(steps)), "trace_union")
        return dict(vertices=b["vertices"], counts=b["counts"], continuation=Particles(b["cont"], ncont.value),
                    endpoints=b.get("endpoints"), seed_steps=b.get("seed_steps"), total_steps=int(steps.value))

    def trace_union_round(self, particles: Particles, region_slots: int, mode: int = 2, parameter: float = 0.75,
                          buffers=None, colors: DeviceBuffer = None):
        """One fused round: union trace (hybrid engine, grouped slots) + prune + index + colour.

        Returns the outputs of the three stages (vertices, indices, colours) besides the trace's."""
        n = particles.n
        b = buffers or self.union_buffers(n, region_slots, endpoints=True)
        vcap = max(1, n * region_slots)
        if "out_vertices" not in b:
            b["out_vertices"] = self.ctx.alloc(vcap * 16)
            b["out_indices"] = self.ctx.alloc(vcap * 4)
            b["out_colors"] = self.ctx.alloc(vcap * 16)
        col = colors if colors is not None else b["out_colors"]
        ncont, nv, ni, steps = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
        self.ctx.check(lib().pa_gpu_tracer_trace_union_round(
            self._h, C.c_void_p(particles.ptr), C.c_size_t(n), C.c_size_t(region_slots),
            C.c_void_p(b["vertices"].ptr), C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.byref(ncont),
            C.c_int(mode), C.c_float(parameter), C.c_void_p(b["out_vertices"].ptr), C.c_size_t(vcap), C.byref(nv),
            C.c_void_p(b["out_indices"].ptr), C.c_size_t(vcap), C.byref(ni), C.c_void_p(col.ptr),
            C.c_void_p(b["endpoints"].ptr), C.c_void_p(b["seed_steps"].ptr), C.byref(steps)), "trace_union_round")
        return dict(slots=b["vertices"], counts=b["counts"], continuation=Particles(b["cont"], ncont.value),
                    vertices=b["out_vertices"], n_vertices=nv.value, indices=b["out_indices"], n_indices=ni.value,
                    colors=col, endpoints=b["endpoints"], seed_steps=b["seed_steps"], total_steps=int(steps.value))

    def trace_union_round_async(self, particles: Particles, region_slots: int, buffers, result: DeviceBuffer,
                                mode: int = 2, parameter: float = 0.75):
        """The fused round without its read-back (pa_gpu_tracer_trace_union_round_async): the
        ROUND_RESULT_DTYPE record lands in `result` on the device; buffers as trace_union_round's."""
        n = particles.n
        b = buffers
        vcap = max(1, n * region_slots)
        self.ctx.check(lib().pa_gpu_tracer_trace_union_round_async(
            self._h, C.c_void_p(particles.ptr), C.c_size_t(n), C.c_size_t(region_slots),
            C.c_void_p(b["vertices"].ptr), C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.c_int(mode),
            C.c_float(parameter), C.c_void_p(b["out_vertices"].ptr), C.c_size_t(vcap),
            C.c_void_p(b["out_indices"].ptr), C.c_size_t(vcap), C.c_void_p(b["out_colors"].ptr),
            C.c_void_p(b["endpoints"].ptr), C.c_void_p(b["seed_steps"].ptr), C.c_void_p(result.ptr)),
            "trace_union_round_async")

    def trace_union_round_pipelined(self, particles: Particles, region_slots: int, buffers, result: DeviceBuffer,
                                    mode: int = 2, parameter: float = 0.75, vertex_capacity: Optional[int] = None):
        """The asynchronous round whose finish runs on the context's second stream while the next
        round traces (pa_gpu_tracer_trace_union_round_pipelined).  Alternate two sets of buffers
        and results; Context.round_join / synchronize order later work after the finishes."""
        n = particles.n
        b = buffers
        vcap = vertex_capacity or max(1, n * region_slots)
        self.ctx.check(lib().pa_gpu_tracer_trace_union_round_pipelined(
            self._h, C.c_void_p(particles.ptr), C.c_size_t(n), C.c_size_t(region_slots),
            C.c_void_p(b["vertices"].ptr), C.c_void_p(b["counts"].ptr), C.c_void_p(b["cont"].ptr), C.c_int(mode),
            C.c_float(parameter), C.c_void_p(b["out_vertices"].ptr), C.c_size_t(vcap),
            C.c_void_p(b["out_indices"].ptr), C.c_size_t(vcap), C.c_void_p(b["out_colors"].ptr),
            C.c_void_p(b["endpoints"].ptr), C.c_void_p(b["seed_steps"].ptr), C.c_void_p(result.ptr)),
            "trace_union_round_pipelined")

    def round_buffers(self, n: int, region_slots: int):
        """Device buffers of one fused round (slots, counts, outputs)."""
        b = self.union_buffers(n, region_slots, endpoints=True)
        vcap = max(1, n * region_slots)
        b["out_vertices"] = self.ctx.alloc(vcap * 16)
        b["out_indices"] = self.ctx.alloc(vcap * 4)
        b["out_colors"] = self.ctx.alloc(vcap * 16)
        return b

    def union_buffers(self, n: int, region_slots: int, endpoints: bool = False):
        n1 = max(n, 1)
        n8 = (n1 + 7) // 8 * 8  # the grouped order rounds the seeds up to groups of 8
        b = dict(vertices=self.ctx.alloc(n8 * region_slots * 16), counts=self.ctx.alloc(n1 * 4),
                 cont=self.ctx.alloc(n1 * 48))
        if endpoints:
            b["endpoints"] = self.ctx.alloc(n1 * 16)
            b["seed_steps"] = self.ctx.alloc(n1 * 4)
        return b

    def advect(self, particles: Particles):
        """particle_advector::advect, one round in place (particle_advector.cpp:61-141)."""
        state = self.ctx.alloc(max(particles.n, 1) * 4)
        slot = self.ctx.alloc(max(particles.n, 1) * 4)
        steps = C.c_uint64()
        self.ctx.check(lib().pa_gpu_advect(self._h, C.c_void_p(particles.ptr), C.c_size_t(particles.n),
                                           C.c_void_p(state.ptr), C.c_void_p(slot.ptr), C.byref(steps)), "advect")
        return state.download(np.int32, particles.n), slot.download(np.int32, particles.n), int(steps.value)

    @staticmethod
    def prune(ctx: Context, vertices: DeviceBuffer, n: int):
        """particle_tracer::prune of one chunk (particle_tracer.cpp:402-409)."""
        out = ctx.alloc(max(n, 1) * 16)
        m = C.c_size_t()
        ctx.check(lib().pa_gpu_prune(ctx.handle, C.c_void_p(vertices.ptr), C.c_size_t(n), C.c_void_p(out.ptr),
                                     C.byref(m)), "prune")
        return out, m.value

    @staticmethod
    def prune_counted(ctx: Context, slots: int, stride_slots: int, counts: DeviceBuffer, n: int, capacity=None):
        """Pack each particle's counted slots (no invalid fill needed)."""
        capacity = capacity or max(1, n * stride_slots)
        out = ctx.alloc(capacity * 16)
        m = C.c_size_t()
        offs = ctx.alloc((n + 1) * 8)
        ctx.check(lib().pa_gpu_prune_counted(ctx.handle, C.c_void_p(slots), C.c_size_t(stride_slots),
                                             C.c_void_p(counts.ptr), C.c_size_t(n), C.c_void_p(out.ptr),
                                             C.c_size_t(capacity), C.byref(m), C.c_void_p(offs.ptr)), "prune_counted")
        return out, m.value, offs

    @staticmethod
    def prune_counted_group8(ctx: Context, slots: int, region_slots: int, counts: DeviceBuffer, n: int,
                             capacity: int, out: DeviceBuffer = None, offs: DeviceBuffer = None):
        """Pack 8-seed grouped union slots (hybrid engine) into the pruned particle-major layout."""
        out = ctx.alloc(max(capacity, 1) * 16) if out is None else out
        offs = ctx.alloc((n + 1) * 8) if offs is None else offs
        m = C.c_size_t()
        ctx.check(lib().pa_gpu_prune_counted_group8(
            ctx.handle, C.c_void_p(slots), C.c_size_t(region_slots), C.c_void_p(counts.ptr), C.c_size_t(n),
            C.c_void_p(out.ptr), C.c_size_t(capacity), C.byref(m), C.c_void_p(offs.ptr)), "prune_counted_group8")
        return out, m.value, offs

    @staticmethod
    def prune_counted_chunked(ctx: Context, slots: int, region_slots: int, counts: DeviceBuffer, n: int,
                              capacity: int, out: DeviceBuffer = None, offs: DeviceBuffer = None):
        """Pack chunked union slots (quad engine) into the pruned particle-major layout."""
        out = ctx.alloc(max(capacity, 1) * 16) if out is None else out
        offs = ctx.alloc((n + 1) * 8) if offs is None else offs
        m = C.c_size_t()
        ctx.check(lib().pa_gpu_prune_counted_chunked(
            ctx.handle, C.c_void_p(slots), C.c_size_t(region_slots), C.c_void_p(counts.ptr), C.c_size_t(n),
            C.c_void_p(out.ptr), C.c_size_t(capacity), C.byref(m), C.c_void_p(offs.ptr)), "prune_counted_chunked")
        return out, m.value, offs

    @staticmethod
    def prune_counted_step_major(ctx: Context, slots: int, counts: DeviceBuffer, n: int, capacity: int,
                                 out: DeviceBuffer = None, offs: DeviceBuffer = None):
        """Pack step-major union slots (row stride n) into the pruned particle-major layout."""
        out = ctx.alloc(max(capacity, 1) * 16) if out is None else out
        offs = ctx.alloc((n + 1) * 8) if offs is None else offs
        m = C.c_size_t()
        ctx.check(lib().pa_gpu_prune_counted_step_major(
            ctx.handle, C.c_void_p(slots), C.c_size_t(n), C.c_void_p(counts.ptr), C.c_void_p(out.ptr),
            C.c_size_t(capacity), C.byref(m), C.c_void_p(offs.ptr)), "prune_counted_step_major")
        return out, m.value, offs


POOL_ACTIVE, POOL_DONE, POOL_INVALID = 0, 1, 2


class _PoolC(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in ("position", "iteration", "remaining", "steps", "status", "field",
                                                 "request", "request_done")] + [("n_requests", C.c_int32)]


class ParticlePool:
    """Particle pool over up to 16 resident fields (pa_gpu_pool_*): batched seed sets and the
    serving loop's step engine.  Each entry follows a one-block reference trace on its field."""

    def __init__(self, ctx: Context, capacity: int, requests: int = 0):
        self.ctx, self.capacity = ctx, capacity
        c = max(capacity, 1)
        self.position = ctx.alloc(c * 16)
        self.iteration = ctx.alloc(c * 4)
        self.remaining = ctx.alloc(c * 4)
        self.steps = ctx.alloc(c * 4)
        self.status = ctx.alloc(c * 4)
        self.field = ctx.alloc(c * 4)
        self.requests = int(requests)
        self.request = ctx.alloc(c * 4) if requests else None
        self.request_done = ctx.upload(np.zeros(requests, np.int32)) if requests else None
        self.status.upload(np.full(c, POOL_INVALID, np.int32))  # entries never initialised are skipped
        self.n = 0

    def _c(self, offset: int = 0):
        def at(b, size):
            return C.c_void_p(b.ptr + offset * size) if b is not None else None
        return _PoolC(at(self.position, 16), at(self.iteration, 4), at(self.remaining, 4), at(self.steps, 4),
                      at(self.status, 4), at(self.field, 4), at(self.request, 4),
                      C.c_void_p(self.request_done.ptr) if self.request_done is not None else None, self.requests)

    def init(self, particles: "Particles", n_fields: int, offset: int = 0):
        """Entries [offset, offset + particles.n) from particles (field = vector_field_index,
        request = original_rank)."""
        assert offset + particles.n <= self.capacity
        pc = self._c(offset)
        st = lib().pa_gpu_pool_init(self.ctx.handle, C.c_void_p(particles.ptr), C.c_size_t(particles.n),
                                    C.c_int32(n_fields), C.byref(pc))
        self.n = max(self.n, offset + particles.n)  # rejected entries are initialised as invalid
        self.ctx.check(st, "pool_init")

    def run(self, tracer: "particle_tracer", fields, n: int = None, max_iterations: int = 0, waves_per_cu: int = 0):
        arr = (C.c_void_p * len(fields))(*[f.handle for f in fields])
        steps = C.c_uint64()
        pc = self._c()
        self.ctx.check(lib().pa_gpu_pool_run(tracer._h, arr, C.c_int32(len(fields)), C.byref(pc),
                                             C.c_size_t(self.n if n is None else n), C.c_int32(max_iterations),
                                             C.c_int32(waves_per_cu), C.byref(steps)), "pool_run")
        return int(steps.value)

    def run_vertices(self, tracer: "particle_tracer", fields, slot_base: DeviceBuffer, slots: DeviceBuffer,
                     counts: DeviceBuffer, n: int = None, max_iterations: int = 0, waves_per_cu: int = 0):
        """run() that also writes each entry's one-block curve slots (pa_gpu_pool_run_vertices):
        slot_base int64 per entry, slots vec4, counts int32 per entry."""
        arr = (C.c_void_p * len(fields))(*[f.handle for f in fields])
        steps = C.c_uint64()
        pc = self._c()
        self.ctx.check(lib().pa_gpu_pool_run_vertices(
            tracer._h, arr, C.c_int32(len(fields)), C.byref(pc), C.c_size_t(self.n if n is None else n),
            C.c_int32(max_iterations), C.c_int32(waves_per_cu), C.c_void_p(slot_base.ptr), C.c_void_p(slots.ptr),
            C.c_size_t(slots.nbytes // 16), C.c_void_p(counts.ptr), C.byref(steps)), "pool_run_vertices")
        return int(steps.value)

    def results(self, n: int = None):
        n = self.n if n is None else n
        return dict(position=self.position.download(np.float32, 4 * n).reshape(n, 4),
                    steps=self.steps.download(np.int32, n), status=self.status.download(np.int32, n))


class index_generator:
    """pa::index_generator (index_generator.cpp:10-24)."""

    @staticmethod
    def generate(ctx: Context, vertices: DeviceBuffer, n: int):
        out = ctx.alloc(max(n, 1) * 4)
        m = C.c_size_t()
        ctx.check(lib().pa_gpu_index_generate(ctx.handle, C.c_void_p(vertices.ptr), C.c_size_t(n),
                                              C.c_void_p(out.ptr), C.byref(m)), "index_generate")
        return out, m.value


class color_generator:
    """pa::color_generator (color_generator.hpp:10-32; color_generator.cpp:16-98)."""

    hsl_constant_s, hsl_constant_l, hsv_constant_s, hsv_constant_v, rgb = range(5)
    magnitude = 100
    MODES = {"hsl_constant_s": 0, "hsl_constant_l": 1, "hsv_constant_s": 2, "hsv_constant_v": 3, "rgb": 4,
             "magnitude": 100}

    @staticmethod
    def generate(ctx: Context, vertices: DeviceBuffer, n_vertices: int, indices: DeviceBuffer, n_indices: int,
                 mode=4, constant_parameter: float = 0.5, colors: Optional[DeviceBuffer] = None):
        if isinstance(mode, str):
            mode = color_generator.MODES[mode]
        if colors is None:
            colors = ctx.upload(np.zeros((max(n_vertices, 1), 4), np.float32))
        ctx.check(lib().pa_gpu_color_generate(ctx.handle, C.c_void_p(vertices.ptr), C.c_size_t(n_vertices),
                                              C.c_void_p(indices.ptr), C.c_size_t(n_indices), C.c_int(mode),
                                              C.c_float(constant_parameter), C.c_void_p(colors.ptr)),
                  "color_generate")
        return colors


class flow_map_generator:
    """pa::flow_map_generator (flow_map_generator.hpp:12-28) for one block without neighbours."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.tracer = particle_tracer(ctx)
        self.tracer.set_integrator(EULER)  # particle_advector.hpp:50 default

    def set_vector_field(self, field: VectorField):
        self.field = field
        self.tracer.set_local_vector_field(field)

    def set_integrator(self, integrator):
        self.tracer.set_integrator(integrator)

    def set_step_size(self, step_size: float):
        self.tracer.set_step_size(step_size)

    def generate(self, block_size, iterations: int, resolution_scale: float = 1.0, rank: int = 0):
        dims = [int(np.float32(b) * np.float32(resolution_scale)) for b in block_size]
        out = self.ctx.alloc(max(1, dims[0] * dims[1] * dims[2]) * 12)
        d = (C.c_int32 * 3)()
        steps = C.c_uint64()
        self.ctx.check(lib().pa_gpu_flow_map_generate(self.tracer.handle, _i3(block_size), C.c_uint64(iterations),
                                                      C.c_float(resolution_scale), C.c_int32(rank),
                                                      C.c_void_p(out.ptr), d, C.byref(steps)), "flow_map")
        assert list(d) == dims
        spacing = self.field.spacing / np.float32(resolution_scale)
        return out, tuple(dims), spacing, int(steps.value)


class ftle_map_generator:
    """pa::ftle_map_generator (ftle_map_generator.hpp:13-40), textbook FP64 form."""

    regular, fractional_anisotropy = 0, 1

    def __init__(self, ctx: Context, mode: int = 0):
        self.ctx = ctx
        self.mode = mode
        self.time = 1.0

    def set_flow_map(self, flow_map: DeviceBuffer, dims, spacing):
        self.flow_map, self.dims, self.spacing = flow_map, tuple(dims), np.asarray(spacing, np.float32)

    def set_time(self, time: float):
        self.time = float(time)

    def generate(self, want_lambda=False):
        n = self.dims[0] * self.dims[1] * self.dims[2]
        out = self.ctx.alloc(max(n, 1) * 4)
        lm = self.ctx.alloc(max(n, 1) * 8) if want_lambda else None
        self.ctx.check(lib().pa_gpu_ftle_generate(self.ctx.handle, C.c_void_p(self.flow_map.ptr), _i3(self.dims),
                                                  _f3(self.spacing), C.c_double(self.time), C.c_int(self.mode),
                                                  C.c_void_p(out.ptr), C.c_void_p(lm.ptr) if lm else None), "ftle")
        return out, lm


# ----------------------------------------------------------------- hand-overs
# particle::serialize field order (pa/include/pa/math/particle.hpp:16-28): the
# 40-byte records boost::mpi sends.
WIRE_DTYPE = np.dtype([("position", "<f4", (4,)), ("remaining_iterations", "<i4"), ("vector_field_index", "<i4"),
                       ("original_rank", "<i4"), ("original_voxel", "<i4", (3,))])
assert WIRE_DTYPE.itemsize == 40


class _CurveChunk(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("colors", C.c_void_p), ("indices", C.c_void_p), ("n_vertices", C.c_size_t),
                ("n_indices", C.c_size_t), ("first_particle", C.c_size_t), ("n_particles", C.c_size_t)]


def round_chunks(ctx: Context, counts: DeviceBuffer, n: int, region_slots: int, vertices: DeviceBuffer, n_vertices: int,
                 indices: DeviceBuffer, n_indices: int, colors: DeviceBuffer = None, max_vertices_per_chunk: int = 0,
                 max_chunks: int = 4096):
    """A fused round's outputs as the renderer's chunks (pa_gpu_round_chunks; ray_tracer.cpp:156-212):
    a list of dicts with device pointers and sizes; indices are rewritten relative to each chunk."""
    arr = (_CurveChunk * max(1, max_chunks))()
    nc = C.c_size_t()
    ctx.check(lib().pa_gpu_round_chunks(
        ctx.handle, C.c_void_p(counts.ptr), C.c_size_t(n), C.c_size_t(region_slots), C.c_size_t(max_vertices_per_chunk),
        C.c_void_p(vertices.ptr), C.c_size_t(n_vertices), C.c_void_p(indices.ptr if indices is not None else None),
        C.c_size_t(n_indices), C.c_void_p(colors.ptr if colors is not None else None), arr, C.c_size_t(max_chunks),
        C.byref(nc)), "round_chunks")
    return [dict(vertices=arr[c].vertices, colors=arr[c].colors, indices=arr[c].indices,
                 n_vertices=arr[c].n_vertices, n_indices=arr[c].n_indices, first_particle=arr[c].first_particle,
                 n_particles=arr[c].n_particles) for c in range(nc.value)]


def particles_to_wire(ctx: Context, particles: "Particles") -> DeviceBuffer:
    """Device particles as serialize records (WIRE_DTYPE), on the device."""
    out = ctx.alloc(max(particles.n, 1) * 40)
    ctx.check(lib().pa_gpu_particles_to_wire(ctx.handle, C.c_void_p(particles.ptr), C.c_size_t(particles.n),
                                             C.c_void_p(out.ptr)), "particles_to_wire")
    return out


def particles_from_wire(ctx: Context, wire: DeviceBuffer, n: int) -> "Particles":
    out = ctx.alloc(max(n, 1) * 48)
    ctx.check(lib().pa_gpu_particles_from_wire(ctx.handle, C.c_void_p(wire.ptr), C.c_size_t(n), C.c_void_p(out.ptr)),
              "particles_from_wire")
    return Particles(out, n)


def load_balance_take(ctx: Context, particles: "Particles", surplus_counts):
    """load_balance_distribute's surplus lists (particle_tracer.cpp:168-184) as serialize records:
    returns (wire buffer, offsets[7], particles kept)."""
    counts = (C.c_size_t * 6)(*[int(c) for c in surplus_counts])
    offs = (C.c_size_t * 7)()
    keep = C.c_size_t()
    total = sum(int(c) for c in surplus_counts)
    wire = ctx.alloc(max(total, 1) * 40)
    ctx.check(lib().pa_gpu_load_balance_take(ctx.handle, C.c_void_p(particles.ptr), C.c_size_t(particles.n), counts,
                                             C.c_void_p(wire.ptr), offs, C.byref(keep)), "load_balance_take")
    return wire, list(offs), keep.value


def read_vector_block(path: str, offset, ghosted, block, want_data: bool = True):
    """The host read of data_io::load_vector_field (data_io.cpp:91-115), no GPU needed:
    returns (xyz [g0][g1][g2][3] float32, offset, size, spacing) as FP32 arrays."""
    g = [int(x) for x in ghosted]
    xyz = np.empty((g[0], g[1], g[2], 3), np.float32) if want_data else np.empty((2, 2, 2, 3), np.float32)
    gg = g if want_data else [2, 2, 2]
    off, size, sp = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
    err = C.create_string_buffer(512)
    st = lib().pa_gpu_read_vector_block(path.encode(), (C.c_int64 * 3)(*[int(x) for x in offset]), _i3(gg),
                                        _i3(block), xyz.ctypes.data_as(C.c_void_p), off, size, sp, err,
                                        C.c_size_t(512))
    if st != OK:
        raise PaError(st, "read_vector_block: " + err.value.decode())
    return xyz, np.array(off[:], np.float32), np.array(size[:], np.float32), np.array(sp[:], np.float32)


def read_vector_dims(path: str):
    """data_io::load_dimensions (data_io.cpp:36-40): node counts of the HDF5 dataset "vectors"."""
    dims = (C.c_int64 * 3)()
    err = C.create_string_buffer(512)
    st = lib().pa_gpu_read_vector_dims(path.encode(), dims, err, C.c_size_t(512))
    if st != OK:
        raise PaError(st, "read_vector_dims: " + err.value.decode())
    return [int(d) for d in dims]


ROUND_RESULT_DTYPE = np.dtype([("total_steps", "<u8"), ("n_continuation", "<u4"), ("flags", "<u4"),
                               ("reserved_", "<u8", (2,)), ("n_vertices", "<i8"), ("n_indices", "<i8")])
assert ROUND_RESULT_DTYPE.itemsize == 48


class Graph:
    """A captured round (pa_gpu_graph_*): the asynchronous calls made on ctx inside
    `with Graph.capture(ctx) as g:` are replayed by g.launch()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._h = C.c_void_p()

    @classmethod
    def capture(cls, ctx: Context):
        return _Capture(cls(ctx))

    def launch(self):
        self.ctx.check(lib().pa_gpu_graph_launch(self.ctx.handle, self._h), "graph_launch")

    def close(self):
        if self._h and self._h.value:
            lib().pa_gpu_graph_destroy(self._h)
            self._h = C.c_void_p()


class _Capture:
    def __init__(self, g: Graph):
        self.g = g

    def __enter__(self):
        self.g.ctx.check(lib().pa_gpu_graph_begin(self.g.ctx.handle), "graph_begin")
        return self.g

    def __exit__(self, exc_type, exc, tb):
        st = lib().pa_gpu_graph_end(self.g.ctx.handle, C.byref(self.g._h))
        if exc_type is None:
            self.g.ctx.check(st, "graph_end")
        return False


def pool_init_async(ctx: Context, pool: "ParticlePool", particles: "Particles", n_fields: int,
                    rejected: DeviceBuffer, zero_requests: bool = True):
    pc = pool._c()
    ctx.check(lib().pa_gpu_pool_init_async(ctx.handle, C.c_void_p(particles.ptr), C.c_size_t(particles.n),
                                           C.c_int32(n_fields), C.byref(pc), C.c_int32(int(zero_requests)),
                                           C.c_void_p(rejected.ptr)), "pool_init_async")
    pool.n = max(pool.n, particles.n)


def pool_run_async(pool: "ParticlePool", tracer: "particle_tracer", fields, result: DeviceBuffer, n: int = None,
                   max_iterations: int = 0, waves_per_cu: int = 0):
    arr = (C.c_void_p * len(fields))(*[f.handle for f in fields])
    pc = pool._c()
    pool.ctx.check(lib().pa_gpu_pool_run_async(tracer._h, arr, C.c_int32(len(fields)), C.byref(pc),
                                               C.c_size_t(pool.n if n is None else n), C.c_int32(max_iterations),
                                               C.c_int32(waves_per_cu), C.c_void_p(result.ptr)), "pool_run_async")


class DeviceServer:
    """The device-resident serving loop (pa_gpu_serve_*): a persistent kernel takes each submitted
    request's particles as soon as they are published; results land in pinned host memory."""

    def __init__(self, ctx: Context, tracer: "particle_tracer", fields, particle_capacity: int,
                 request_slots: int = 1024, waves_per_cu: int = 0, idle_polls: int = 0):
        self.ctx = ctx
        self._tracer = tracer
        self._fields = list(fields)
        arr = (C.c_void_p * len(fields))(*[f.handle for f in fields])
        self._h = C.c_void_p()
        ctx.check(lib().pa_gpu_serve_create(tracer._h, arr, C.c_int32(len(fields)), C.c_size_t(particle_capacity),
                                            C.c_int32(request_slots), C.c_int32(waves_per_cu), C.c_int32(idle_polls),
                                            C.byref(self._h)), "serve_create")

    def submit(self, particles: np.ndarray):
        """Returns the request slot, or None when the ring or the slots are full."""
        p = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        slot = C.c_int32()
        st = lib().pa_gpu_serve_submit(self._h, p.ctypes.data_as(C.c_void_p), C.c_size_t(len(p)), C.byref(slot))
        if st == E_CAPACITY and "larger than the ring" not in lib().pa_gpu_last_error(self.ctx.handle).decode():
            return None  # ring or slots full for now: try again once requests are collected
        self.ctx.check(st, "serve_submit")
        return slot.value

    def done(self, slot: int) -> bool:
        d = C.c_int32()
        self.ctx.check(lib().pa_gpu_serve_poll(self._h, C.c_int32(slot), C.byref(d)), "serve_poll")
        return bool(d.value)

    def collect(self, slot: int, n: int, want_endpoints: bool = True):
        ends = np.empty((n, 4), np.float32) if want_endpoints else None
        steps = np.empty(n, np.int32)
        total = C.c_uint64()
        self.ctx.check(lib().pa_gpu_serve_collect(self._h, C.c_int32(slot),
                                                  ends.ctypes.data_as(C.c_void_p) if ends is not None else None,
                                                  steps.ctypes.data_as(C.c_void_p), C.byref(total)), "serve_collect")
        return ends, steps, int(total.value)

    def stop(self):
        steps, launches = C.c_uint64(), C.c_uint64()
        self.ctx.check(lib().pa_gpu_serve_stop(self._h, C.byref(steps), C.byref(launches)), "serve_stop")
        return int(steps.value), int(launches.value)

    def close(self):
        if self._h and self._h.value:
            lib().pa_gpu_serve_destroy(self._h)
            self._h = C.c_void_p()
