"""Velocity restriction of the forward wavelet transform -- TEST
INFRASTRUCTURE ONLY (a numpy restatement, the checker of the HIP kernel).

Restrict_velo (wavelet.f90:1154-1174): on a uniform hierarchy every fine
edge is active, and the coarse edge of parent node (i, j) is the mean of the
two fine edges that bisect it, in the lozenge frame of the child (2i, 2j):
  RT: (RT(child) + RT(child + E)) / 2
  DG: (DG(child + NE) + DG(child)) / 2
  UP: (UP(child) + UP(child + N)) / 2
Owned coarse nodes only (the poles own no edges); halos come from the
coarse level's ghost fill.
"""
import numpy as np


def restrict_velo(fine_layout, coarse_layout, uf):
    """uf [L][3][NC_fine] -> [L][3][NC_coarse] on the owned coarse edges
    (zero elsewhere)."""
    nc, Sc, Sf = coarse_layout.n, coarse_layout.S, fine_layout.S
    L = uf.shape[0]
    out = np.zeros((L, 3, coarse_layout.ncells))
    lz, j, i = np.meshgrid(np.arange(10), np.arange(nc), np.arange(nc), indexing="ij")
    p = (lz * Sc * Sc + (j + 2) * Sc + (i + 2)).ravel()
    c = (lz * Sf * Sf + (2 * j + 2) * Sf + (2 * i + 2)).ravel()
    out[:, 0, p] = 0.5 * (uf[:, 0, c] + uf[:, 0, c + 1])
    out[:, 1, p] = 0.5 * (uf[:, 1, c + Sf + 1] + uf[:, 1, c])
    out[:, 2, p] = 0.5 * (uf[:, 2, c] + uf[:, 2, c + Sf])
    return out
