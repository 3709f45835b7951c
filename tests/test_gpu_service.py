"""The service adapter end to end on the GPU: settings bytes over the PAIR
channel in, image bytes out (pars/source/service.cpp:37-67), with
pars::pipeline::execute's reuse rules (pipeline.cpp:20-50): a camera change
renders again without tracing, an advection change traces again, a dataset
change loads again.  The traced curves are the oracle's, and the image is
the point splat of those curves (checked against a host restatement of the
projection on all but a few boundary pixels)."""
import os

import numpy as np
import pytest

import pars_amd as pa
from pars_amd import service as S

pytestmark = pytest.mark.gpu

H5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors_small.h5")


def _settings(**kw):
    base = dict(mode="streamlines", dataset_filepath="abc:40", seed_generation_stride=[3, 3, 3],
                seed_generation_iterations=80, particle_tracing_integrator="runge_kutta_4",
                particle_tracing_step_size=0.02, color_generation_mode="hsv_constant_s",
                color_generation_free_parameter=0.6, raytracing_camera_position=[3.0, 3.0, -6.0],
                raytracing_camera_forward=[0.0, 0.0, 1.0], raytracing_camera_up=[0.0, 1.0, 0.0],
                raytracing_image_size=[96, 64], raytracing_iterations=1)
    base.update(kw)
    return S.SETTINGS(**base)


def _splat(verts, cols, s):
    """Host restatement of pa_gpu_render_points (nearest vertex per pixel)."""
    w, h = s.raytracing_image_size
    pos = np.array(s.raytracing_camera_position, np.float64)
    f = np.array(s.raytracing_camera_forward, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross(f, np.array(s.raytracing_camera_up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    t = np.tan(np.radians(34.0))
    keep = ~np.all(verts == [-2, -2, -2, 0], axis=1) & ~np.all(verts == [-1, -1, -1, 0], axis=1)
    d = verts[keep, :3].astype(np.float64) - pos
    c = cols[keep]
    z = d @ f
    ok = z > 1e-6
    x = (d @ r)[ok] / (z[ok] * t * w / h)
    y = (d @ u)[ok] / (z[ok] * t)
    px = (x * 0.5 + 0.5) * w
    py = (y * 0.5 + 0.5) * h
    inside = (px >= 0) & (px < w) & (py >= 0) & (py < h)
    img = np.zeros((h, w), bool)
    img[py[inside].astype(int), px[inside].astype(int)] = True
    return img


def test_service_requests(ctx, orc):
    ch = S.PairChannel()
    pipe = S.Pipeline(ctx)
    first = _settings()
    ch.client.send(first.SerializeToString())
    assert S.run(ch.service, pipe, max_requests=1) == 1
    img = S.IMAGE()
    img.ParseFromString(ch.client.recv(timeout=5))
    w, h = first.raytracing_image_size
    assert list(img.size) == [w, h] and len(img.data) == 4 * w * h
    rgba = np.frombuffer(img.data, np.uint8).reshape(h, w, 4)
    lit = rgba[..., 3] > 0
    assert lit.sum() > 100
    # the curves of that request are the oracle's (one block, same stride and R)
    r = pipe.round
    nv = r["n_vertices"]
    verts = r["vertices"].download(np.float32, 4 * nv).reshape(nv, 4)
    cols = r["colors"].download(np.float32, 4 * nv).reshape(nv, 4)
    f = pipe.field
    O = orc.Field(f.read(), f.offset, f.size, f.spacing)
    seeds = orc.seed_generate(f.offset, f.size, f.spacing * np.float32(3), 80, 0)
    ou = orc.trace_union(seeds, O, [None] * 6, [-1] * 6, [0] * 6, dt=0.02, integrator=orc.RK4)
    assert verts.tobytes() == ou["vertices"].tobytes()
    # the image is the splat of those vertices
    want = _splat(verts, cols, first)
    assert np.count_nonzero(want != lit) <= max(3, lit.size // 500)
    # camera change: no new trace, a different image
    moved = _settings(raytracing_camera_position=[3.0, 2.0, -7.0])
    res = pipe.execute(moved)
    assert not res.traced and res.image.data != img.data
    # advection change: traced again, more vertices with longer curves
    longer = _settings(seed_generation_iterations=120)
    res2 = pipe.execute(longer)
    assert res2.traced and res2.vertices > nv
    # unsupported integrator: an empty image, the service keeps running
    bad = _settings(particle_tracing_integrator="runge_kutta_dormand_prince_5")
    empty = S.IMAGE()
    empty.ParseFromString(pipe.handle(bad.SerializeToString()))
    assert list(empty.size) == [] and empty.data == b""
    # dataset from the HDF5 fixture (data_io's read of the whole domain as one rank)
    res3 = pipe.execute(_settings(dataset_filepath=H5, seed_generation_stride=[2, 2, 2],
                                  seed_generation_iterations=30, particle_tracing_step_size=0.05,
                                  raytracing_camera_position=[0.6, 0.4, -2.0]))
    assert res3.traced and res3.vertices > 0 and pipe.field.dims == (20, 18, 16)
    pipe.close()
