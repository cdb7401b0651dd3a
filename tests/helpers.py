"""Shared set-up for the parity tests: one seeded load fed to both codes."""
import numpy as np

import oracle as O
from swap_amd import physics

# tiles of the sort in swap_amd/csrc/k_sort.hip (TILE_X, TILE_Y, TILE_Z of
# swap_amd/csrc/swap_internal.h) and its chunk of consecutive particles
TILE = (4, 8, 8)
SORT_CHUNK = 4096


def unpack(cell):
    cell = np.asarray(cell, np.uint32)
    return (cell & 2047).astype(np.int64), ((cell >> 11) & 1023).astype(np.int64), \
        (cell >> 21).astype(np.int64)


def tiles(cell, nc):
    """tile index and cell index inside the tile of packed cells, and the
    number of tiles"""
    i, j, k = unpack(cell)
    ntx = (nc[0] + TILE[0] - 1) // TILE[0]
    nty = (nc[1] + TILE[1] - 1) // TILE[1]
    ntz = (nc[2] + TILE[2] - 1) // TILE[2]
    tile = i // TILE[0] + ntx * (j // TILE[1] + nty * (k // TILE[2]))
    local = i % TILE[0] + TILE[0] * (j % TILE[1] + TILE[1] * (k % TILE[2]))
    return tile, local, ntx * nty * ntz


def sort_perm(cell, nc):
    """the permutation of the particle sort: stable by (tile, chunk of
    SORT_CHUNK consecutive source positions, cell in tile), through the
    oracle's stable counting sort of that composite key"""
    tile, local, nt = tiles(cell, nc)
    nch = (len(cell) + SORT_CHUNK - 1) // SORT_CHUNK
    cells = TILE[0] * TILE[1] * TILE[2]
    chunk = np.arange(len(cell), dtype=np.int64) // SORT_CHUNK
    key = (tile * nch + chunk) * cells + local
    nkeys = nt * nch * cells
    assert nkeys < 2 ** 32
    return O.stable_sort(key.astype(np.uint32), nkeys)


def rel(a, b):
    """max-norm error of a relative to the max-norm of b"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


class Pair:
    """A GPU context and an FP64 oracle on the same grid and the same load."""

    def __init__(self, ncx=16, ncy=8, ncz=8, ppc=8, precision=32, sort_every=2, flags=4,
                 seed=20261203, width=64, load=True, capacity=None, vsw=400.0):
        from swap_amd import Swap  # noqa: F401
        self.p = p = physics.workload(ncx, ncy, ncz, ppc, vsw=vsw)
        self.s = physics.make_swap(p, capacity=capacity, precision=precision,
                                   sort_every=sort_every, flags=flags, seed=seed)
        self.o = O.Oracle(ncx + 2, ncy + 2, ncz + 2, p["dx"], p["dx"], p["dx"], p["dt"],
                          p["ntsub"], p["alpha"], p["beta"], width=width)
        self.n = p["npart"]
        self.nx, self.ny, self.nz = ncx + 2, ncy + 2, ncz + 2
        dt = self.o.dtype
        self.st = {
            "b1": self.o.grid(), "b12": self.o.grid(), "b1p2": self.o.grid(),
            "bt": self.o.grid(), "btc": self.o.grid(), "np": self.o.grid(1),
            "up": self.o.grid(), "aj": self.o.grid(), "E": self.o.grid(),
            "nu": np.full((self.nz, self.ny, self.nx), p["nu"], dt),
            "vplus": self.o.parts(self.n), "vminus": self.o.parts(self.n),
            "Ep": self.o.parts(self.n), "wght": self.o.parts(self.n, 8),
            "ijkp": np.zeros((3, self.n), np.int32), "Ni_tot": self.n,
        }
        for b in ("b1", "b12", "b1p2", "bt"):
            self.st[b][2] = p["b0z"]
        if load:
            self.s.load_maxwellian(self.n, p["vsw"], p["vth"])
            self.pull_particles()

    def pull_particles(self):
        dt = self.o.dtype
        self.st["xp"] = np.ascontiguousarray(self.s.xp().astype(dt))
        self.st["vp"] = np.ascontiguousarray(self.s.vp().astype(dt))
        self.st["vp1"] = np.ascontiguousarray(self.s.vp1().astype(dt))

    def push_fields(self, names=("b1", "b12")):
        for nm in names:
            self.s.set_field(nm, self.st[nm])
