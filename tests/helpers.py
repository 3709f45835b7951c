"""Shared set-up of the parity tests: the same blocks on the device (through
the C ABI) and in the oracle (on the decoded node values)."""
import numpy as np

import pars_amd as pa
from pars_amd import workloads as W

f32 = np.float32


class ShareFields:
    """Own block + 6 neighbours of one rank, on the device and in the oracle."""

    def __init__(self, ctx, orc, domain_n, ranks, rank, storage=pa.FP32, field="abc", modes=None,
                 layout=pa.LAYOUT_NODES):
        self.share = W.share(domain_n, ranks, rank)
        sh = self.share
        infos = [sh.local] + list(sh.neighbors)
        self.dev = []
        self.orc = []
        self.f64 = []
        for info in infos:
            if info is None:
                self.dev.append(None)
                self.orc.append(None)
                self.f64.append(None)
                continue
            u64 = W.abc_block_f64(info, sh.h) if field == "abc" else W.fourier_block_f64(info, sh.h, modes)
            d = pa.VectorField.from_host(ctx, u64.astype(np.float32), info.offset, info.size, info.spacing, storage)
            d.set_layout(layout)
            self.dev.append(d)
            self.orc.append(orc.Field(d.read(), info.offset, info.size, info.spacing))
            self.f64.append(orc.Field64(u64, info.offset.astype(np.float64), info.size.astype(np.float64),
                                        [sh.h] * 3))

    @property
    def local(self):
        return self.dev[0]

    @property
    def neighbors(self):
        return self.dev[1:]

    def tracer(self, ctx, dt, integrator=pa.RUNGE_KUTTA_4, interp=pa.INTERP_REFERENCE, engine=pa.ENGINE_COOPERATIVE,
               waves=0):
        t = pa.particle_tracer(ctx)
        t.set_engine(engine, waves)
        t.set_local_vector_field(self.local)
        t.set_neighbor_vector_fields(self.neighbors, self.share.neighbor_ranks)
        t.set_neighbor_face_masks(self.share.face_masks)
        t.set_integrator(integrator)
        t.set_step_size(dt)
        t.set_interpolation(interp)
        return t

    def oracle_kwargs(self, dt, integrator, interp):
        return dict(neighbors=self.orc[1:], neighbor_rank=self.share.neighbor_ranks, dt=dt, integrator=integrator,
                    interp=interp)


def single_block(ctx, orc, n, storage=pa.FP32, h=None, layout=pa.LAYOUT_NODES):
    """ABC on n^3 nodes over [0, 2pi], one block without neighbours."""
    h = h if h is not None else 2 * np.pi / (n - 1)
    info = W.BlockInfo(0, (0, 0, 0), (0, 0, 0), (n, n, n))
    u64 = W.abc_block_f64(info, h)
    hf = f32(h)
    off = np.zeros(3, np.float32)
    size = np.array([f32(n) * hf] * 3, np.float32)  # data_io: size = block * spacing
    sp = np.array([hf] * 3, np.float32)
    d = pa.VectorField.from_host(ctx, u64.astype(np.float32), off, size, sp, storage)
    d.set_layout(layout)
    o = orc.Field(d.read(), off, size, sp)
    o64 = orc.Field64(u64, off.astype(np.float64), size.astype(np.float64), [h] * 3)
    return d, o, o64


def oracle_integrator(orc, integ):
    return {pa.EULER: orc.EULER, pa.RUNGE_KUTTA_4: orc.RK4, pa.CLASSIC_RK4: orc.CLASSIC_RK4}[integ]
