"""Host-side bounds of the hybrid tracer's two computed addresses, restated
from the kernel's arithmetic and checked on the CPU:

- vertex stores: a wave's raw buffer stores are offsets from its first
  slot (group8_slot of the wave's first seed, slot 0) in a descriptor whose
  range ends at the drop offset; every store of every lane and slot must end
  at or below it for region_slots up to the engine's limit, and the limit is
  tight (one more slot reaches the drop offset);
- the cell-record gather: the record index of a lane, with the cell clamped
  into [0, dims - 2] per axis (also for positions outside the block, or not
  a number), stays below the block's cell count.

The constants are read from the kernel sources, so a change there is
checked against the same arithmetic."""
import os
import re

import numpy as np

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pars_amd", "csrc")


def _match(fname, pattern):
    with open(os.path.join(CSRC, fname)) as fh:
        m = re.search(pattern, fh.read())
    assert m, (fname, pattern)
    return m


K_DROP = int(_match("pa_device.hpp", r"constexpr int kDrop = (0x[0-9A-Fa-f]+);").group(1), 16)
MAX_REGION = (1 << int(_match("trace_quad.hip",
                              r"constexpr int64_t kHybridMaxRegion = \(int64_t\(1\) << (\d+)\) - 1;").group(1))) - 1


def group8_slot(p, s, region):
    return ((p >> 3) * region + s) * 8 + (p & 7)


def max_store_end(region, waves=(0, 1, 7)):
    """Largest end (bytes) of a 16-B store of any lane of the given waves, any slot, relative to
    the wave's base."""
    worst = 0
    for w in waves:
        wave0 = 64 * w
        base = group8_slot(wave0, 0, region)
        lanes = np.arange(64, dtype=np.int64) + wave0
        for s in (0, region - 1):  # offsets grow with s: the last slot is the largest
            off = (group8_slot(lanes, s, region) - base) * 16 + 16
            worst = max(worst, int(off.max()))
    return worst


def test_store_offsets_closed_form():
    """The closed form the kernel asserts (lane 63 = seed 7 of group 7, slot region - 1) is the
    largest store end of a wave, for every region (brute force over all slots on small ones)."""
    for region in (1, 2, 3, 8, 37, 1002):
        brute = 0
        lanes = np.arange(64, dtype=np.int64)
        for s in range(region):
            brute = max(brute, int(((group8_slot(lanes, s, region) - group8_slot(0, 0, region)) * 16 + 16).max()))
        closed = ((7 * region + (region - 1)) * 8 + 7) * 16 + 16
        assert brute == closed == max_store_end(region)


def test_store_offsets_within_the_descriptor():
    assert MAX_REGION == (1 << 21) - 1
    assert max_store_end(MAX_REGION) <= K_DROP
    assert max_store_end(MAX_REGION + 1) > K_DROP  # the limit is tight
    # the headline workload's region (R + extra slots) is far inside
    assert max_store_end(1002 + 40) < K_DROP // 100


def test_clamped_record_index_in_range():
    rng = np.random.default_rng(11)
    for dims in [(2, 2, 2), (17, 17, 17), (129, 129, 129), (5, 300, 9)]:
        d = np.array(dims, np.int64)
        n_cells = int(np.prod(d - 1))
        cstride0, cstride1 = (d[1] - 1) * (d[2] - 1), d[2] - 1
        q = rng.uniform(-3 * d.max(), 3 * d.max(), (20000, 3)).astype(np.float32)
        q[:50] = np.nan
        q[50:100] = np.inf
        q[100:150] = -np.inf
        q[150:200] = (d - 1).astype(np.float32)  # on the far faces
        # the kernel: fminf(fmaxf(floorf(q), 0), dims - 2); fmaxf(NaN, 0) is 0
        fl = np.floor(q)
        fl = np.where(np.isnan(fl), np.float32(0), fl)
        m = np.minimum(np.maximum(fl, 0), (d - 2).astype(np.float32)).astype(np.int64)
        ci = m[:, 0] * cstride0 + m[:, 1] * cstride1 + m[:, 2]
        assert ci.min() >= 0 and ci.max() < n_cells
        assert ci.max() < 2 ** 32  # the kernel forms the index in 32 bits


def logical_block(b, nb, remap):
    """The XCD-aware workgroup order of the tracers (pa_device.hpp), restated."""
    if not remap:
        return b
    if remap >= 2:
        c = remap
        whole = nb // (8 * c) * (8 * c)
        if b >= whole:
            return b
        x, k = b % 8, b // 8
        return ((k // c) * 8 + x) * c + k % c
    q, r = nb // 8, nb % 8
    x, k = b % 8, b // 8
    return x * q + (x if x < r else r) + k


def test_xcd_orders_are_bijections():
    """Every workgroup order maps the grid onto itself one to one (each particle traced once), and a run
    order gives every XCD (b % 8) runs of consecutive logical workgroups from every part of the list."""
    for nb in (1, 7, 8, 9, 63, 64, 100, 1024, 1031, 4097):
        for remap in (0, 1, 2, 3, 4, 8, 16, 64):
            m = [logical_block(b, nb, remap) for b in range(nb)]
            assert sorted(m) == list(range(nb)), (nb, remap)
    nb, c = 1024, 4
    for x in range(8):
        mine = sorted(logical_block(b, nb, c) for b in range(x, nb, 8))
        runs = [mine[i:i + c] for i in range(0, len(mine), c)]
        assert all(r == list(range(r[0], r[0] + c)) for r in runs)
        assert mine[0] < nb // 8 and mine[-1] >= nb - nb // 8  # spread over the whole list
