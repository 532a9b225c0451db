"""Shared setup of parity tests: grids, fixture states, oracle."""
import numpy as np

import wavetrisk_hydrostatic_amd as W
from oracle.oracle import Oracle, geometry_of


def grid_of_fixture(f):
    return W.Grid(level=int(f["level"]), min_level=int(f["min_level"]), domain_level=int(f["domain_level"]),
                  radius=float(f["radius"]), omega=float(f["omega"]))


def fixture_state(f, grid):
    """Initial state of a trisk fixture on the layout: owned cells, pole
    masters and frozen halo slots from the reference, other halo cells 0
    (the ghost fill provides them)."""
    nc = grid.layout.ncells
    node_src = grid.array("node_src")
    edge_src = grid.array("edge_src").reshape(nc, 3).T
    mass, temp, velo = (np.nan_to_num(f[k].copy()) for k in ("mass0", "temp0", "velo0"))
    lz, i, j = grid.layout.coords()
    owned = (i >= 0) & (i < grid.layout.n) & (j >= 0) & (j < grid.layout.n)
    halo_fill = (~owned) & (node_src >= 0) & (node_src != np.arange(nc))
    mass[:, halo_fill] = 0.0
    temp[:, halo_fill] = 0.0
    for e in range(3):
        fill = (~owned) & (edge_src[e] >= 0)
        velo[:, e, fill] = 0.0
    topo = np.nan_to_num(f["topo"].copy())
    return topo, mass, temp, velo


def oracle_of(grid, topo, f=None, grav=9.80616, rho0=1028.0):
    if f is not None:
        grav, rho0 = float(f["grav"]), float(f["ref_density"])
    return Oracle(grid.layout, geometry_of(grid), topo, grav=grav, rho0=rho0)
