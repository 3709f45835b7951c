"""Synthetic workloads of the advection path: block geometry of a rank's
share (the partitioner and data_io arithmetic), the ABC flow, seeded
random-Fourier curl fields, and the seeded generators they use.

Geometry follows pa/source/stages/partitioner.cpp:8-53 (prime-factor rank
grid, ghost layer +1 on the positive side unless at the domain border,
neighbour order -x,+x,-y,+y,-z,+z) and pa/source/stages/data_io.cpp:113-114
(offset = float(int offset) * spacing, size = float(block) * spacing, FP32).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

ABC = (math.sqrt(3.0), math.sqrt(2.0), 1.0)
DEFAULT_SEED = 20261203


def prime_factors(n: int) -> List[int]:
    """pa/include/pa/math/prime_factorize.hpp: ascending prime factors."""
    out, p = [], 2
    while n > 1:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    return out


def rank_grid(domain: Tuple[int, int, int], ranks: int) -> Tuple[int, int, int]:
    """partitioner::set_domain_size (partitioner.cpp:24-45)."""
    factors = prime_factors(ranks)
    cur = list(domain)
    grid = [1, 1, 1]
    while factors:
        d = 0
        for i in range(3):
            if cur[i] > cur[d]:
                d = i
        f = factors.pop()
        cur[d] //= f
        grid[d] *= f
    return tuple(grid)


def unravel(index: int, dims) -> Tuple[int, int, int]:
    out = [0, 0, 0]
    for i in (2, 1, 0):
        out[i] = index % dims[i]
        index //= dims[i]
    return tuple(out)


def ravel(multi, dims) -> int:
    r = 0
    for i in range(3):
        r = r * dims[i] + multi[i]
    return r


@dataclass
class BlockInfo:
    rank: int
    multi_rank: Tuple[int, int, int]
    int_offset: Tuple[int, int, int]
    ghosted: Tuple[int, int, int]
    offset: np.ndarray = field(default=None)
    size: np.ndarray = field(default=None)
    spacing: np.ndarray = field(default=None)
    neighbor_ranks: Tuple[int, ...] = ()
    face_mask: int = 0


@dataclass
class Share:
    """One rank's share: its block and the six face neighbours."""

    domain: Tuple[int, int, int]
    grid: Tuple[int, int, int]
    block: Tuple[int, int, int]
    h: float                       # physical spacing (double)
    local: BlockInfo
    neighbors: List[Optional[BlockInfo]]

    @property
    def neighbor_ranks(self):
        return [nb.rank if nb is not None else -1 for nb in self.neighbors]

    @property
    def face_masks(self):
        return [nb.face_mask if nb is not None else 0 for nb in self.neighbors]


def _block(rank, grid, block, h32):
    mr = unravel(rank, grid)
    off = tuple(block[i] * mr[i] for i in range(3))
    gh = tuple(block[i] + (1 if mr[i] + 1 < grid[i] else 0) for i in range(3))
    nbr = []
    mask = 0
    for k, (axis, step) in enumerate([(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)]):
        m = list(mr)
        m[axis] += step
        if 0 <= m[axis] < grid[axis]:
            nbr.append(ravel(m, grid))
            mask |= 1 << k
        else:
            nbr.append(-1)
    sp = np.array([h32] * 3, np.float32)
    offset = np.array([np.float32(off[i]) * sp[i] for i in range(3)], np.float32)
    size = np.array([np.float32(block[i]) * sp[i] for i in range(3)], np.float32)
    return BlockInfo(rank, mr, off, gh, offset, size, sp, tuple(nbr), mask)


def share(domain_n: int, ranks: int, rank: int) -> Share:
    """Geometry of `rank` in a cube domain of domain_n^3 nodes over [0, 2pi]."""
    domain = (domain_n,) * 3
    grid = rank_grid(domain, ranks)
    block = tuple(domain[i] // grid[i] for i in range(3))
    h = 2.0 * math.pi / (domain_n - 1)
    h32 = np.float32(h)
    local = _block(rank, grid, block, h32)
    neighbors = [(_block(r, grid, block, h32) if r >= 0 else None) for r in local.neighbor_ranks]
    return Share(domain, grid, block, h, local, neighbors)


def abc_block_f64(info: BlockInfo, h: float, A=ABC[0], B=ABC[1], C=ABC[2]) -> np.ndarray:
    """ABC flow at global nodes (int_offset + ijk) * h, FP64, [d0][d1][d2][3]."""
    ax = [(info.int_offset[i] + np.arange(info.ghosted[i], dtype=np.float64)) * h for i in range(3)]
    x, y, z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    out = np.empty(info.ghosted + (3,), np.float64)
    out[..., 0] = A * np.sin(z) + C * np.cos(y)
    out[..., 1] = B * np.sin(x) + A * np.cos(z)
    out[..., 2] = C * np.sin(y) + B * np.cos(x)
    return out


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood, OOPSLA 2014); uniform doubles in [0, 1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self) -> float:
        """Box-Muller, one value per call (the second is discarded)."""
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def fourier_modes(n_modes: int, seed: int, kmax: int = 8) -> np.ndarray:
    """Random vector potential A = sum a_m sin(k_m.x + phi_m), k_m uniform in
    1..kmax per axis, a_m ~ N(0,1)^3 / |k_m|^2; returns rows
    (k0, k1, k2, c0, c1, c2, phi) with c = k x a, so u = curl A =
    sum cos(k.x + phi) c (divergence-free by construction)."""
    rng = SplitMix64(seed)
    rows = []
    for _ in range(n_modes):
        k = np.array([rng.randint(1, kmax) for _ in range(3)], np.float64)
        a = np.array([rng.normal() for _ in range(3)], np.float64) / float(k @ k)
        phi = 2.0 * math.pi * rng.uniform()
        c = np.cross(k, a)
        rows.append([k[0], k[1], k[2], c[0], c[1], c[2], phi])
    return np.array(rows, np.float64)


def fourier_block_f64(info: BlockInfo, h: float, modes: np.ndarray) -> np.ndarray:
    ax = [(info.int_offset[i] + np.arange(info.ghosted[i], dtype=np.float64)) * h for i in range(3)]
    x, y, z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    out = np.zeros(info.ghosted + (3,), np.float64)
    for m in modes:
        cs = np.cos(m[0] * x + m[1] * y + m[2] * z + m[6])
        out[..., 0] += cs * m[3]
        out[..., 1] += cs * m[4]
        out[..., 2] += cs * m[5]
    return out


def quantize(xyz: np.ndarray, storage: int) -> np.ndarray:
    """FP32 values as they read back from BF16 / FP16 storage (round to nearest even)."""
    a = np.ascontiguousarray(xyz, np.float32)
    if storage == 0:
        return a.copy()
    if storage == 2:
        return a.astype(np.float16).astype(np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)
