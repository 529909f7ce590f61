"""Seeded synthetic particle/grid workloads for the transfer benchmarks.

Random numbers come from a counter-based SplitMix64 stream (vectorised in
numpy): u = (splitmix64(seed + (i+1)*golden) >> 11) * 2^-53, normals by
Box-Muller.  Every field of every workload is a pure function of its seed,
so tests and benchmarks on any machine see identical inputs.

Workloads (unit cube, uniform mesh of n^3 cells, h = 1/n, two ghost cells):
  apic_uniform(n, ppc)   every cell holds ppc particles, v ~ N(0,1),
                         B ~ N(0,0.1), m = h^3/ppc   (32^3 x 8 ppc is the
                         CPU-runnable plumbing case)
  apic_sphere(n, ppc)    cells whose centre lies within 0.3 of the cube
                         centre hold ppc particles; v = (0,0,1) x (x - c)
                         + N(0,0.05), B ~ N(0,0.1)   (256^3 x 8 ppc, ~15.2M
                         particles, is the headline workload)
  apic_cube(n, ppc)      the central cube [1/4, 3/4]^3, F = I + N(0,0.01)
  polypic_uniform(n,ppc) 27 velocity modes per component ~ N(0,0.1),
                         mode 0 ~ N(0,1)
Particles are shuffled (a random permutation from the same stream) so a
cell sort has real work to do.
"""
from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
HALO = 2


def _splitmix(seed, start, count):
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + 1 + count, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential draws from one seeded counter-based stream."""

    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def bits(self, n):
        z = _splitmix(self.seed, self.pos, n)
        self.pos += n
        return z

    def uniform(self, n):
        return (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n, std=1.0):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))  # 1 - u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return z * std

    def permutation(self, n):
        return np.argsort(self.bits(n), kind="stable")


@dataclass
class Workload:
    name: str
    n: int                 # cells per dimension
    ppc: int
    h: float
    low: np.ndarray        # local mesh low corner (ghosts included)
    x: np.ndarray          # [P, 3] float64
    m: np.ndarray          # [P]
    c: np.ndarray = None   # [P, 4, 3] APIC state
    F: np.ndarray = None   # [P, 3, 3]
    modes: np.ndarray = None  # [P, 27, 3]
    extra: dict = field(default_factory=dict)

    @property
    def np(self):
        return self.x.shape[0]

    @property
    def node_window(self):
        """(lo, n) of the node window that covers every stencil up to cubic."""
        nn = self.n + 1 + 2 * HALO
        return np.array([0, 0, 0], np.int32), np.array([nn, nn, nn], np.int32)

    @property
    def cell_region(self):
        """(cell_lo, ncell) of the cells particles may occupy."""
        return np.array([HALO] * 3, np.int32), np.array([self.n] * 3, np.int32)


def _fill_cells(st, cells, ppc, h):
    """ppc uniform positions in each listed cell (cells [C,3] int)."""
    C = cells.shape[0]
    u = st.uniform(C * ppc * 3).reshape(C * ppc, 3)
    return (np.repeat(cells, ppc, axis=0) + u) * h


def _finish(st, x, ppc, h):
    P = x.shape[0]
    perm = st.permutation(P)
    return x[perm]


def apic_uniform(n=32, ppc=8, seed=0x5EED0000):
    st = Stream(seed)
    h = 1.0 / n
    ii = np.arange(n)
    cells = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    x = _finish(st, _fill_cells(st, cells, ppc, h), ppc, h)
    P = x.shape[0]
    c = np.empty((P, 4, 3))
    c[:, 0, :] = st.normal(P * 3).reshape(P, 3)
    c[:, 1:, :] = st.normal(P * 9, 0.1).reshape(P, 3, 3)
    m = np.full(P, h ** 3 / ppc)
    return Workload(f"apic_uniform_{n}^3x{ppc}", n, ppc, h, np.full(3, -HALO * h), x, m, c=c)


def sphere_cells(n, radius=0.3):
    h = 1.0 / n
    ctr = (np.arange(n) + 0.5) * h - 0.5
    d2 = ctr[:, None, None] ** 2 + ctr[None, :, None] ** 2 + ctr[None, None, :] ** 2
    return np.argwhere(d2 <= radius * radius)


def apic_sphere(n=256, ppc=8, seed=0x5EED0001, radius=0.3):
    st = Stream(seed)
    h = 1.0 / n
    x = _finish(st, _fill_cells(st, sphere_cells(n, radius), ppc, h), ppc, h)
    P = x.shape[0]
    c = np.empty((P, 4, 3))
    # rigid rotation omega = (0,0,1) about the centre, plus N(0, 0.05)
    c[:, 0, 0] = -(x[:, 1] - 0.5)
    c[:, 0, 1] = x[:, 0] - 0.5
    c[:, 0, 2] = 0.0
    c[:, 0, :] += st.normal(P * 3, 0.05).reshape(P, 3)
    c[:, 1:, :] = st.normal(P * 9, 0.1).reshape(P, 3, 3)
    m = np.full(P, h ** 3 / ppc)
    return Workload(f"apic_sphere_{n}^3x{ppc}", n, ppc, h, np.full(3, -HALO * h), x, m, c=c)


def apic_cube(n=256, ppc=8, seed=0x5EED0004):
    st = Stream(seed)
    h = 1.0 / n
    ii = np.arange(n // 4, 3 * n // 4)
    cells = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    x = _finish(st, _fill_cells(st, cells, ppc, h), ppc, h)
    P = x.shape[0]
    c = np.empty((P, 4, 3))
    c[:, 0, :] = st.normal(P * 3).reshape(P, 3)
    c[:, 1:, :] = st.normal(P * 9, 0.1).reshape(P, 3, 3)
    F = np.eye(3)[None] + st.normal(P * 9, 0.01).reshape(P, 3, 3)
    m = np.full(P, h ** 3 / ppc)
    return Workload(f"apic_cube_{n}^3x{ppc}", n, ppc, h, np.full(3, -HALO * h), x, m, c=c, F=F)


def polypic_uniform(n=128, ppc=8, seed=0x5EED0002, nmode=27):
    st = Stream(seed)
    h = 1.0 / n
    ii = np.arange(n)
    cells = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    x = _finish(st, _fill_cells(st, cells, ppc, h), ppc, h)
    P = x.shape[0]
    modes = st.normal(P * nmode * 3, 0.1).reshape(P, nmode, 3)
    modes[:, 0, :] = st.normal(P * 3).reshape(P, 3)
    m = np.full(P, h ** 3 / ppc)
    return Workload(f"polypic_uniform_{n}^3x{ppc}", n, ppc, h, np.full(3, -HALO * h), x, m,
                    modes=modes)


def touched_nodes(x, low, h, order=2):
    """Exact number of distinct nodes inside any particle's stencil."""
    t = (x - low) / h
    if order == 2:
        base = np.floor(t + 0.5).astype(np.int64) - 1
    elif order == 3:
        base = np.floor(t).astype(np.int64) - 1
    else:
        base = np.floor(t).astype(np.int64)
    nk = order + 1
    span = base.max(0) + nk + 1
    mark = np.zeros(tuple(span), bool)
    for i in range(nk):
        for j in range(nk):
            for k in range(nk):
                mark[base[:, 0] + i, base[:, 1] + j, base[:, 2] + k] = True
    return int(mark.sum())
