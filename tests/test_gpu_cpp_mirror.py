"""The C++ stage mirror (pars_amd/include/pa_amd/stages.hpp) drives the
GPU through the C ABI like pars' pipeline drives pa::; its outputs match the
oracle's seed -> trace -> prune -> index -> colour chain."""
import os
import subprocess

import numpy as np
import pytest

import pars_amd as pa
from pars_amd import workloads as W

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "stages_demo")
H5 = os.path.join(os.path.dirname(__file__), "golden", "vectors_small.h5")


def test_cpp_mirror_pipeline(orc, tmp_path):
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    n, R, dt = 40, 150, 0.01
    h = 2 * np.pi / (n - 1)
    u = W.abc_block_f64(W.BlockInfo(0, (0, 0, 0), (0, 0, 0), (n, n, n)), h).astype(np.float32)
    fin = tmp_path / "field.f32"
    u.tofile(fin)
    out = str(tmp_path / "run")
    r = subprocess.run([EXE, str(fin), str(n), str(R), repr(dt), out, H5], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    hf = np.float32(h)
    size = [np.float32(n) * hf] * 3
    seeds = orc.seed_generate((0, 0, 0), size, [np.float32(3.0) * hf] * 3, R, 0)
    got_seeds = np.fromfile(out + ".seeds", orc.PARTICLE_DTYPE)
    assert got_seeds.tobytes() == seeds.tobytes()
    F = orc.Field(u, (0, 0, 0), size, [hf] * 3)
    _, chunks, _, _, _ = orc.trace(seeds, F, dt=np.float32(dt), integrator=orc.RK4)
    v = orc.prune(chunks[0])
    got_v = np.fromfile(out + ".vertices", np.float32).reshape(-1, 4)
    assert got_v.tobytes() == v.tobytes()
    idx = orc.index_generate(v)
    assert np.array_equal(np.fromfile(out + ".indices", np.int32), idx)
    col = orc.color_generate(v, idx, 2, 0.75)
    got_c = np.fromfile(out + ".colors", np.float32).reshape(-1, 4)
    np.testing.assert_allclose(got_c, col, atol=1e-6)
    # the fused round on one block produces the stage chain's curves bit for bit
    assert np.fromfile(out + ".round_vertices", np.float32).tobytes() == got_v.tobytes()
    assert np.array_equal(np.fromfile(out + ".round_indices", np.int32), idx)
    # colours are compared where the colour stage writes them (the others
    # are undefined in the reference: colors.resize leaves them uninitialised)
    rc = np.fromfile(out + ".round_colors", np.float32).reshape(-1, 4)
    written = np.zeros(len(v), bool)
    written[idx] = True
    ends = np.ones(len(idx), bool)
    ends[:-1] = idx[1:] != idx[:-1] + 1
    ends[-1] = False
    written[idx[ends] + 1] = True
    assert rc[written].tobytes() == got_c[written].tobytes()
    # the pool ends every seed where the round ends it
    ends_round = np.fromfile(out + ".round_endpoints", np.float32).reshape(-1, 4)
    pool_pos = np.fromfile(out + ".pool_positions", np.float32).reshape(-1, 4)
    assert pool_pos.tobytes() == ends_round.tobytes()
    # two rounds in flight through round_pipeline: the first equals the synchronous round, the second
    # (the first half of the seeds) is the prefix of its curves
    assert np.fromfile(out + ".pipe1_vertices", np.float32).tobytes() == got_v.tobytes()
    p2 = np.fromfile(out + ".pipe2_vertices", np.float32).reshape(-1, 4)
    assert p2.tobytes() == got_v[:len(p2)].tobytes()
    assert np.array_equal(np.fromfile(out + ".pipe2_indices", np.int32), orc.index_generate(p2))
    # the resident server ends every request's particles where the pool does
    served = np.fromfile(out + ".served_endpoints", np.float32).reshape(-1, 4)
    assert served.tobytes() == pool_pos.tobytes()
    # wire records: the first 40 bytes of each particle, in particle.hpp's field order; back unchanged
    wire = np.fromfile(out + ".wire", np.uint8).reshape(-1, 40)
    assert wire.tobytes() == seeds.view(np.uint8).reshape(-1, 48)[:, :40].tobytes()
    back = np.fromfile(out + ".from_wire", orc.PARTICLE_DTYPE)
    assert back.tobytes() == seeds.tobytes()
    # the HDF5 block as data_io loads it (same numbers as the host read of the fixture)
    xyz, off, size_, sp = pa.read_vector_block(H5, (2, 1, 3), (9, 8, 7), (8, 7, 6))
    info = np.fromfile(out + ".h5_info", np.float32)
    assert np.array_equal(info[:3], [9, 8, 7])
    assert info[3:6].tobytes() == np.asarray(off, np.float32).tobytes()
    assert info[6:9].tobytes() == np.asarray(size_, np.float32).tobytes()
    assert info[9:12].tobytes() == np.asarray(sp, np.float32).tobytes()
