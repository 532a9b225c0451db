"""Synthetic initial states on the lozenge-tile layout (host-side numpy).

Every state is defined by analytic fields of longitude/latitude evaluated at
the coordinates of each cell, halo cells included, and edge velocities are
the projection of a (u_zonal, v_merid) field on the edge direction at the
edge midpoint, as the reference's test cases do (sphere.f90:315-341
proj_vel; e.g. test/DCMIP2008c5/test_case_module.f90:565-592).  Values of
frozen halo slots (never ghost filled) therefore match what the reference
would hold there.  Edges whose far end lies outside the tile are left 0,
like the reference's unset ghost edges.

Scalars of the Boussinesq layer model: mass = rho0*dz, temp = rho0*dz*(1 -
rho/rho0) (cal_density, ops.f90:1383-1404); layer 0 is the bottom layer.
"""
from __future__ import annotations

import numpy as np

from .layout import Layout


def lonlat(x, radius):
    lat = np.arcsin(np.clip(x[2] / radius, -1.0, 1.0))
    lon = np.arctan2(x[1], x[0])
    return lon, lat


def _edge_endpoints(layout: Layout):
    """(tail, head, valid) cell indices of the 3 edges of every cell."""
    S, nc = layout.S, layout.ncells
    c = np.arange(nc)
    lz, i, j = layout.coords()
    hi = layout.n + layout.halo
    tails = [c, c + S + 1, c]
    heads = [c + 1, c, c + S]
    valid = [(i + 1 < hi), (i + 1 < hi) & (j + 1 < hi), (j + 1 < hi)]
    out = []
    for e in range(3):
        t = np.where(valid[e], tails[e], c)
        h = np.where(valid[e], heads[e], c)
        out.append((t, h, valid[e]))
    return out


def project_velocity(layout: Layout, x: np.ndarray, radius: float, vel_fun) -> np.ndarray:
    """Velocity along each primal edge (tail -> head) at its midpoint, [3, nc]."""
    out = np.zeros((3, layout.ncells))
    for e, (t, h, ok) in enumerate(_edge_endpoints(layout)):
        p, q = x[:, t], x[:, h]
        mid = p + q
        nrm = np.linalg.norm(mid, axis=0)
        nrm[nrm == 0] = 1.0
        mid = mid * (radius / nrm)
        lon, lat = lonlat(mid, radius)
        ez = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])
        em = np.stack([-np.cos(lon) * np.sin(lat), -np.sin(lon) * np.sin(lat), np.cos(lat)])
        u, v = vel_fun(lon, lat)
        vel = u * ez + v * em
        d = q - p
        dn = np.linalg.norm(d, axis=0)
        good = ok & (dn >= radius * 1e-13)
        dn[~good] = 1.0
        out[e] = np.where(good, np.sum(d / dn * vel, axis=0), 0.0)
    return out


def tc5(grid, nlayers: int = 1, u0: float = 20.0, h0: float = 5960.0, b0: float = 2000.0,
        half_width: float = 1.5e6, rho0: float = 1028.0):
    """Williamson test case 5 (zonal flow over an isolated mountain) as a
    single Boussinesq layer: solid-body zonal wind u0, balanced surface around
    h0, Gaussian mountain of amplitude b0 and half-width 1500 km at 30N, 90W."""
    lay = grid.layout
    x = grid.array("x")
    R, Om, g = grid.radius, grid.omega, 9.80616
    lon, lat = lonlat(x, R)
    lonc, latc = -np.pi / 2, np.pi / 6
    cosd = np.sin(latc) * np.sin(lat) + np.cos(latc) * np.cos(lat) * np.cos(lon - lonc)
    dist = R * np.arccos(np.clip(cosd, -1, 1))
    bottom = b0 * np.exp(-(dist / half_width) ** 2)
    surface = h0 - (R * Om * u0 + 0.5 * u0 ** 2) * np.sin(lat) ** 2 / g
    dz = (surface - bottom) / nlayers
    mass = np.repeat((rho0 * dz)[None, :], nlayers, axis=0)
    temp = np.zeros_like(mass)
    velo1 = project_velocity(lay, x, R, lambda lo, la: (u0 * np.cos(la), 0.0 * la))
    velo = np.repeat(velo1[None], nlayers, axis=0)
    return bottom, mass, temp, velo


def ocean(grid, nlayers: int, seed: int = 0, depth: float = 4000.0, rho_top: float = 1024.0,
          rho_bottom: float = 1028.0, bump: float = 0.5, u0: float = 0.1, rho0: float = 1028.0):
    """Multilayer Boussinesq ocean: stably stratified layers (density linear
    from rho_bottom at the bottom to rho_top at the top), depth minus a seeded
    smooth bathymetry, a seeded Gaussian sea-surface bump and a weak seeded
    velocity field (all single-valued at the poles)."""
    lay = grid.layout
    x = grid.array("x")
    R = grid.radius
    rng = np.random.default_rng(seed)
    lon, lat = lonlat(x, R)
    # smooth bathymetry: a few seeded low-order modes, rms ~ 300 m
    bathy = np.zeros(lay.ncells)
    for _ in range(6):
        a, m, l, ph = rng.normal(0, 150), rng.integers(1, 6), rng.integers(1, 6), rng.uniform(0, 2 * np.pi)
        bathy += a * np.cos(lat) * np.cos(m * lon + ph) * np.cos(l * lat)
    bottom = -depth + bathy
    lonb, latb = rng.uniform(-np.pi, np.pi), rng.uniform(-1.0, 1.0)
    cosd = np.sin(latb) * np.sin(lat) + np.cos(latb) * np.cos(lat) * np.cos(lon - lonb)
    r = np.arccos(np.clip(cosd, -1, 1))
    eta = bump * np.exp(-(r / 0.2) ** 2)
    dz = (eta - bottom) / nlayers
    mass = np.empty((nlayers, lay.ncells))
    temp = np.empty((nlayers, lay.ncells))
    for k in range(nlayers):
        rho = rho_bottom if nlayers == 1 else rho_bottom + (rho_top - rho_bottom) * k / (nlayers - 1)
        mass[k] = rho0 * dz
        temp[k] = rho0 * dz * (1.0 - rho / rho0)
    ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
    base = project_velocity(
        lay, x, R,
        lambda lo, la: (u0 * np.cos(la) * (1 + 0.3 * np.sin(3 * lo + ph1) * np.cos(la)),
                        0.3 * u0 * np.cos(la) ** 2 * np.sin(2 * lo + ph2)))
    velo = np.empty((nlayers, 3, lay.ncells))
    for k in range(nlayers):  # weaker flow with height
        velo[k] = (1.0 - 0.5 * k / max(nlayers - 1, 1)) * base
    return bottom, mass, temp, velo
