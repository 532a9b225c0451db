"""Lozenge-tile layout of one uniform icosahedral level (pure host logic).

10 lozenges of n x n nodes, each a tile of S x S cells (S = n + 2*halo);
cell (lz, i, j) = lz*S*S + (j+halo)*S + (i+halo), i, j in [-halo, n+halo).
Owned nodes: i, j in [0, n).  South pole: cell (0, n, 0); north pole:
cell (5, 0, n) (the reference's pole masters, init.f90:547-569).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Layout:
    level: int
    n: int
    S: int
    halo: int = 2

    @property
    def ncells(self) -> int:
        return 10 * self.S * self.S

    def cell(self, lz, i, j):
        return lz * self.S * self.S + (j + self.halo) * self.S + (i + self.halo)

    @property
    def south_pole(self) -> int:
        return int(self.cell(0, self.n, 0))

    @property
    def north_pole(self) -> int:
        return int(self.cell(5, 0, self.n))

    def coords(self):
        """(lz, i, j) arrays of every cell."""
        c = np.arange(self.ncells)
        lz = c // (self.S * self.S)
        r = c % (self.S * self.S)
        return lz, r % self.S - self.halo, r // self.S - self.halo

    def owned_cells(self, with_poles: bool = True) -> np.ndarray:
        """Cells whose values are prognostic: owned nodes (+ the two pole masters)."""
        lz, i, j = self.coords()
        m = (i >= 0) & (i < self.n) & (j >= 0) & (j < self.n)
        cells = np.nonzero(m)[0]
        if with_poles:
            cells = np.concatenate([cells, [self.south_pole, self.north_pole]])
        return cells

    def owned_edge_cells(self) -> np.ndarray:
        """Cells whose three edges are owned (the poles own none)."""
        return self.owned_cells(with_poles=False)

    @property
    def n_dof_per_layer(self) -> int:
        """Prognostic unknowns per layer: 2 scalars per node + 1 per edge."""
        nodes = 10 * self.n * self.n + 2
        edges = 30 * self.n * self.n
        return 2 * nodes + edges
