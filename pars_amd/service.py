"""The pars service's wire adapter over the MI355X path.

pars_service talks to pars over a ZMQ PAIR socket: the client sends one
serialized `pars.settings` protobuf message, the service answers with one
serialized `pars.image` (pars/source/service.cpp:20-69; messages
pars/proto/settings.proto:5-27 and pars/proto/image.proto:5-9).  This module
keeps that contract:

- `SETTINGS` / `IMAGE` are the two message types, built at run time from
  descriptors with the .proto files' field numbers and types (no protoc in
  this image), so the bytes on the wire are the reference's;
- `Pipeline.execute(settings)` mirrors pars::pipeline::execute
  (pars/source/pipeline.cpp:20-330) for the "streamlines" mode on one rank:
  the dataset (an HDF5 file read as data_io does, or a synthetic field named
  `abc:<n>` / `fourier:<n>:<modes>:<seed>`) is loaded into HBM once and kept
  while `dataset_filepath` stays the same; seeds, the fused trace round
  (trace + prune + index + colour) and the renderer chunks run again only
  when an advection parameter changes; the image is rendered again for camera
  or image-size changes.  The image comes from pa_gpu_render_points, a
  point-splat stand-in for OSPRay's streamline ray tracer (out of scope),
  with OSPRay's camera model and RGBA8 framebuffer layout;
- `run(transport, pipeline)` is service.cpp's loop: receive one request,
  answer it, repeat.  `PairChannel` is an in-process PAIR pair; a zmq PAIR
  socket is used instead when pyzmq is importable (`zmq_transport`).
"""
from __future__ import annotations

import ctypes as C
import math
import queue
from dataclasses import dataclass
from typing import Optional

import numpy as np

from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

from . import (FP32, LAYOUT_CELLS, Context, PaError, VectorField, lib, particle_tracer, round_chunks, seed_generator,
               E_UNSUPPORTED)
from . import workloads as W

# ------------------------------------------------------------------ messages
_F = descriptor_pb2.FieldDescriptorProto
# (name, number, type, repeated): settings.proto:5-27
_SETTINGS_FIELDS = [
    ("mode", 1, _F.TYPE_STRING, False),
    ("volume_type", 2, _F.TYPE_STRING, False),
    ("dataset_filepath", 3, _F.TYPE_STRING, False),
    ("seed_generation_stride", 4, _F.TYPE_INT32, True),
    ("seed_generation_iterations", 5, _F.TYPE_INT32, False),
    ("particle_tracing_integrator", 6, _F.TYPE_STRING, False),
    ("particle_tracing_step_size", 7, _F.TYPE_FLOAT, False),
    ("particle_tracing_load_balance", 8, _F.TYPE_BOOL, False),
    ("color_generation_mode", 9, _F.TYPE_STRING, False),
    ("color_generation_free_parameter", 10, _F.TYPE_FLOAT, False),
    ("raytracing_camera_position", 11, _F.TYPE_FLOAT, True),
    ("raytracing_camera_forward", 12, _F.TYPE_FLOAT, True),
    ("raytracing_camera_up", 13, _F.TYPE_FLOAT, True),
    ("raytracing_image_size", 14, _F.TYPE_INT32, True),
    ("raytracing_streamline_radius", 15, _F.TYPE_FLOAT, False),
    ("raytracing_iterations", 16, _F.TYPE_INT32, False),
]
# image.proto:5-9
_IMAGE_FIELDS = [("size", 1, _F.TYPE_INT32, True), ("data", 2, _F.TYPE_BYTES, False)]


def _build_messages():
    pool = descriptor_pool.DescriptorPool()
    for fname, mname, fields in (("settings.proto", "settings", _SETTINGS_FIELDS),
                                 ("image.proto", "image", _IMAGE_FIELDS)):
        fdp = descriptor_pb2.FileDescriptorProto(name=fname, package="pars", syntax="proto3")
        m = fdp.message_type.add(name=mname)
        for name, number, typ, rep in fields:
            m.field.add(name=name, number=number, type=typ,
                        label=_F.LABEL_REPEATED if rep else _F.LABEL_OPTIONAL)
        pool.Add(fdp)
    return (message_factory.GetMessageClass(pool.FindMessageTypeByName("pars.settings")),
            message_factory.GetMessageClass(pool.FindMessageTypeByName("pars.image")))


SETTINGS, IMAGE = _build_messages()

# pipeline.cpp:128-145 and :244-253: the settings' names of integrators and colour modes
INTEGRATOR_NAMES = ("euler", "modified_midpoint", "runge_kutta_4", "runge_kutta_cash_karp_54",
                    "runge_kutta_dormand_prince_5", "runge_kutta_fehlberg_78", "adams_bashforth_2",
                    "adams_bashforth_moulton_2")
COLOR_MODES = {"hsl_constant_s": 0, "hsl_constant_l": 1, "hsv_constant_s": 2, "hsv_constant_v": 3, "rgb": 4}


# -------------------------------------------------------------------- pipeline
@dataclass
class Result:
    image: object            # IMAGE message
    traced: bool             # whether this request traced (advection or dataset changed)
    steps: int = 0
    vertices: int = 0
    indices: int = 0
    chunks: int = 0


class Pipeline:
    """pars::pipeline::execute for the streamlines mode, on one MI355X."""

    def __init__(self, ctx: Context, storage=FP32, layout=LAYOUT_CELLS):
        self.ctx, self.storage, self.layout = ctx, storage, layout
        self.last = None
        self.field: Optional[VectorField] = None
        self.round = None
        self.region = 0
        self.rgba = None

    # 1.x: data_io (pipeline.cpp:56-100)
    def _load(self, path: str):
        if self.field is not None:
            self.field.release()
        if path.startswith("abc:") or path.startswith("fourier:"):
            parts = path.split(":")
            n = int(parts[1])
            h = 2.0 * math.pi / (n - 1)
            hf = np.float32(h)
            size = np.array([np.float32(n) * hf] * 3, np.float32)
            if parts[0] == "abc":
                kind, params = "abc", W.ABC
            else:
                kind, params = "fourier", W.fourier_modes(int(parts[2]), int(parts[3]))
            f = VectorField.synthesize(self.ctx, kind, params, (n, n, n), (0, 0, 0), h, np.zeros(3, np.float32), size,
                                       np.array([hf] * 3, np.float32), self.storage)
        else:
            dims = _dataset_dims(path)
            f = VectorField.load(self.ctx, path, (0, 0, 0), dims, dims, self.storage)  # one rank: the whole domain
        self.field = f.set_layout(self.layout)

    # 2.0-5.0: seeds, trace rounds, prune, index, colour (pipeline.cpp:107-260)
    def _trace(self, s):
        name = s.particle_tracing_integrator
        if name not in ("euler", "runge_kutta_4"):
            raise PaError(E_UNSUPPORTED, "integrator %r has no GPU path" % name)
        if len(s.seed_generation_stride) != 3:
            raise PaError(1, "seed_generation_stride needs 3 values")
        f = self.field
        stride = f.spacing * np.array(list(s.seed_generation_stride), np.float32)
        R = int(s.seed_generation_iterations)
        seeds = seed_generator.generate(self.ctx, f.offset, f.size, stride, R, 0)
        t = particle_tracer(self.ctx)
        t.set_local_vector_field(f)
        t.set_integrator(name)
        t.set_step_size(s.particle_tracing_step_size)
        self.region = max(R, 1) + 2
        mode = COLOR_MODES.get(s.color_generation_mode, COLOR_MODES["rgb"])
        r = t.trace_union_round(seeds, self.region, mode, s.color_generation_free_parameter)
        if r["continuation"].n:
            raise PaError(5, "segment region overflow")
        self.chunks = round_chunks(self.ctx, r["counts"], seeds.n, self.region, r["vertices"], r["n_vertices"],
                                   r["indices"], r["n_indices"], r["colors"])
        self.round = r
        return r

    # 6.x: the renderer (pipeline.cpp:262-330), here a point splat
    def _render(self, s):
        w, h = (int(s.raytracing_image_size[0]), int(s.raytracing_image_size[1])) \
            if len(s.raytracing_image_size) == 2 else (0, 0)
        img = IMAGE()
        if w < 1 or h < 1:
            return img
        if self.rgba is None or self.rgba.nbytes < 4 * w * h:
            self.rgba = self.ctx.alloc(4 * w * h)
        def f3(v, default):
            return (C.c_float * 3)(*(list(v) if len(v) == 3 else default))
        r = self.round
        st = lib().pa_gpu_render_points(
            self.ctx.handle, C.c_void_p(r["vertices"].ptr if r else None), C.c_void_p(r["colors"].ptr if r else None),
            C.c_size_t(r["n_vertices"] if r else 0), f3(s.raytracing_camera_position, [0, 0, -1]),
            f3(s.raytracing_camera_forward, [0, 0, 1]), f3(s.raytracing_camera_up, [0, 1, 0]), C.c_int32(w),
            C.c_int32(h), C.c_void_p(self.rgba.ptr))
        self.ctx.check(st, "render_points")
        img.size.extend([w, h])
        img.data = self.rgba.download(np.uint8, 4 * w * h).tobytes()
        return img

    def execute(self, s) -> Result:
        if "streamlines" not in s.mode:
            raise PaError(E_UNSUPPORTED, "mode %r: only streamlines run on this path" % s.mode)
        last = self.last
        dataset_changed = last is None or last.dataset_filepath != s.dataset_filepath
        advection_changed = last is None or any(getattr(last, k) != getattr(s, k) for k in (
            "seed_generation_stride", "seed_generation_iterations", "particle_tracing_integrator",
            "particle_tracing_step_size", "particle_tracing_load_balance", "color_generation_mode",
            "color_generation_free_parameter", "raytracing_streamline_radius"))
        if dataset_changed:
            self._load(s.dataset_filepath)
        traced = dataset_changed or advection_changed
        if traced:
            self._trace(s)
        img = self._render(s)
        self.last = SETTINGS()
        self.last.CopyFrom(s)
        r = self.round
        return Result(img, traced, r["total_steps"], r["n_vertices"], r["n_indices"], len(self.chunks))

    def handle(self, request: bytes) -> bytes:
        """One PAIR exchange: settings bytes in, image bytes out (service.cpp:43-66).  A request the path
        cannot serve answers with an empty image, as rank 0's serialize does for an empty framebuffer."""
        s = SETTINGS()
        s.ParseFromString(request)
        try:
            return self.execute(s).image.SerializeToString()
        except PaError:
            return IMAGE().SerializeToString()

    def close(self):
        if self.field is not None:
            self.field.release()
            self.field = None


def _dataset_dims(path: str):
    from . import read_vector_dims
    return read_vector_dims(path)


# ------------------------------------------------------------------ transport
class PairChannel:
    """An in-process stand-in for a connected ZMQ PAIR socket pair: two ends,
    each send() delivered to the other end's recv(), in order."""

    def __init__(self):
        self._q = (queue.Queue(), queue.Queue())
        self.service = _End(self._q[0], self._q[1])
        self.client = _End(self._q[1], self._q[0])


class _End:
    def __init__(self, inbox, outbox):
        self._in, self._out = inbox, outbox

    def send(self, data: bytes):
        self._out.put(bytes(data))

    def recv(self, timeout: Optional[float] = None) -> bytes:
        return self._in.get(timeout=timeout)


def zmq_transport(address: str):
    """A bound ZMQ PAIR socket as service.cpp:24-40 has it (needs pyzmq)."""
    import zmq  # noqa: pyzmq is optional

    sock = zmq.Context.instance().socket(zmq.PAIR)
    sock.bind(address)
    return sock


def run(transport, pipeline: Pipeline, max_requests: Optional[int] = None):
    """service.cpp:37-67: answer requests one at a time, in order."""
    served = 0
    while max_requests is None or served < max_requests:
        request = transport.recv()
        if request is None:
            break
        transport.send(pipeline.handle(request))
        served += 1
    return served
