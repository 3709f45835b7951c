"""The calibration kernels that the measurements divide by: the plain copy
(pa_gpu_copy, byte-counter calibration) and the scattered line reads
(pa_gpu_gather_lines, the rate of reads on unrelated lines).  The line
choice is restated here in numpy (the SplitMix64 finaliser of j ^ seed,
scaled to the line count by a 32x32 high product) and the device's XOR of
everything read must equal the XOR computed from the host copy."""
import ctypes as C

import numpy as np
import pytest

import pars_amd as pa

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def mix32(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return (z ^ (z >> np.uint64(31))).astype(np.uint32)


def gather_xor(words, line_bytes, read_bytes, n_reads, seed):
    n_lines = (len(words) * 4) // line_bytes
    j = np.arange(n_reads, dtype=np.uint64)
    line = (mix32(j ^ np.uint64(seed)).astype(np.uint64) * np.uint64(n_lines)) >> np.uint64(32)
    start = line * np.uint64(line_bytes // 4)
    acc = np.uint32(0)
    for w in range(read_bytes // 4):
        acc ^= np.bitwise_xor.reduce(words[(start + np.uint64(w)).astype(np.int64)])
    return int(acc)


@pytest.mark.parametrize("line_bytes,read_bytes", [(64, 48), (128, 16), (128, 64), (48, 48)])
@pytest.mark.parametrize("n_reads", [1, 1000, 300_001])
def test_gather_lines_reads_what_it_names(ctx, line_bytes, read_bytes, n_reads):
    rng = np.random.default_rng(7)
    nbytes = 1 << 20
    words = rng.integers(0, 1 << 32, size=nbytes // 4, dtype=np.uint64).astype(np.uint32)
    src = ctx.upload(words)
    out = ctx.upload(np.zeros(4, np.uint32))
    seed = 0x1234_5678_9ABC
    ctx.check(pa.lib().pa_gpu_gather_lines(ctx.handle, C.c_void_p(src.ptr), C.c_size_t(nbytes), C.c_int32(line_bytes),
                                           C.c_int32(read_bytes), C.c_int64(n_reads), C.c_uint64(seed),
                                           C.c_void_p(out.ptr)), "gather_lines")
    got = int(out.download(np.uint32, 4)[0])
    assert got == gather_xor(words, line_bytes, read_bytes, n_reads, seed)


def test_gather_lines_rejects_bad_shapes(ctx):
    src = ctx.alloc(4096)
    out = ctx.alloc(16)
    for line, read in [(64, 80), (8, 8), (64, 24), (32, 48)]:
        rc = pa.lib().pa_gpu_gather_lines(ctx.handle, C.c_void_p(src.ptr), C.c_size_t(4096), C.c_int32(line),
                                          C.c_int32(read), C.c_int64(10), C.c_uint64(1), C.c_void_p(out.ptr))
        assert rc == pa.E_INVALID, (line, read)


@pytest.mark.parametrize("nbytes", [16, 4096 + 7, (1 << 22) + 16 * 5 + 3])
def test_copy_is_a_copy(ctx, nbytes):
    rng = np.random.default_rng(nbytes)
    data = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
    a = ctx.upload(data)
    b = ctx.upload(np.zeros(nbytes, np.uint8))
    ctx.check(pa.lib().pa_gpu_copy(ctx.handle, C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes)), "copy")
    assert np.array_equal(b.download(np.uint8, nbytes), data)
