"""Learned correction of the free-surface multigrid: problem draws, training
pairs harvested from the FP64 GPU solver, training on the GPU, and the
comparison against plain V-cycles.

The elliptic problem is the implicit free surface of the split-explicit
ocean (elliptic_lo, barotropic_2d.f90:187-253) on a uniform hierarchy:
    theta1 theta2 dt^2 div(g (|topo| + eta_old) grad u) - u = f.
A draw is a seeded ocean bathymetry (depth 4000 m minus smooth modes), a
seeded old free surface and a seeded right-hand side (smooth bumps plus
grid-scale noise).  A training pair is, for one 8x8 patch of the finest
level and one iterate u_k of plain V-cycles from u = 0, the network's input
(the residual r_k = f - L u_k over the patch scaled by its RMS sigma, the
depth and sin(latitude)) and the target (u* - u_k) / sigma over the patch,
u* the solution to 1e-13.

    python -m wavetrisk_hydrostatic_amd.learned   # harvest, train, save
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import (FS_DIVF, FS_DIVF_OLD, FS_ETA, FS_ETA_OLD, FS_RHS, MLP, SOLVE_FMG, SOLVE_LEARNED, SOLVE_VCYCLE,
               Context, Grid, Hierarchy, MgParams)
from .cases import lonlat, ocean

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_J7.npz")


def draw_fields(grids, seed: int):
    """per level (topography, old free surface, right-hand side) of one draw"""
    rng = np.random.default_rng(seed)
    bseed, eseed = (int(v) for v in rng.integers(0, 2 ** 31, 2))
    amps = rng.normal(0.0, 1.0, 4)
    centres = [(rng.uniform(-np.pi, np.pi), rng.uniform(-1.2, 1.2), rng.uniform(0.05, 0.4)) for _ in range(4)]
    noise_seed = int(rng.integers(0, 2 ** 31))
    out = []
    for g in grids:
        topo, _, _, _ = ocean(g, 1, seed=bseed)
        lon, lat = lonlat(g.array("x"), g.radius)
        erng = np.random.default_rng(eseed)
        lo0, la0 = erng.uniform(-np.pi, np.pi), erng.uniform(-1.0, 1.0)
        eta_old = 0.5 * np.exp(-(_dist(lon, lat, lo0, la0) / 0.2) ** 2)
        f = np.zeros(g.layout.ncells)
        for a, (lo, la, w) in zip(amps, centres):
            f += a * np.exp(-(_dist(lon, lat, lo, la) / w) ** 2)
        f += 0.05 * np.random.default_rng(noise_seed + g.layout.n).standard_normal(g.layout.ncells)
        out.append((topo, eta_old, f))
    return out


class Problem:
    """One draw on the uniform hierarchy coarse..fine (contexts + hierarchy);
    seeds: a list of draws makes it a batch, one ensemble member per draw
    (the hierarchy then solves every member's problem in the same launches)."""

    def __init__(self, grids, seed, dt: float = 1800.0):
        self.grids = grids
        seeds = list(seed) if isinstance(seed, (list, tuple)) else [seed]
        self.members = M = len(seeds)
        draws = [draw_fields(grids, s) for s in seeds]
        self.ctxs, self.topos, self.eta_olds, self.rhs = [], [], [], []
        for k, g in enumerate(grids):
            topo = np.stack([d[k][0] for d in draws]) if M > 1 else draws[0][k][0]
            eta_old = np.stack([d[k][1] for d in draws]) if M > 1 else draws[0][k][1]
            f = np.stack([d[k][2] for d in draws]) if M > 1 else draws[0][k][2]
            self.topos.append(topo)
            self.eta_olds.append(eta_old)
            self.rhs.append(f)
            c = Context(g, nlayers=1, members=M)
            c.set_topography(topo)
            c.fs_set_params(dt)
            c.fs_upload(FS_ETA_OLD, eta_old)
            c.fs_upload(FS_RHS, f)
            c.fs_upload(FS_ETA, np.zeros_like(f))
            self.ctxs.append(c)
        self.h = Hierarchy(self.ctxs)
        self.top = self.ctxs[-1]

    def close(self):
        self.h.close()
        for c in self.ctxs:
            c.close()

    def reset(self):
        for c, g in zip(self.ctxs, self.grids):
            shape = (g.layout.ncells,) if self.members == 1 else (self.members, g.layout.ncells)
            c.fs_upload(FS_ETA, np.zeros(shape))

    def solve(self, method, params):
        self.reset()
        t = time.perf_counter()
        it, res = self.h.solve(method, params)
        return it, res, (time.perf_counter() - t) * 1e3


def _dist(lon, lat, lo, la):
    c = np.sin(la) * np.sin(lat) + np.cos(la) * np.cos(lat) * np.cos(lon - lo)
    return np.arccos(np.clip(c, -1.0, 1.0))


def hierarchy_grids(fine: int = 7, levels: int = 3, domain_level: int = 2):
    return [Grid(level=l, min_level=fine - levels + 1, domain_level=domain_level)
            for l in range(fine - levels + 1, fine + 1)]


def vcycle_params(tol=1e-9, max_vcycle=30):
    """V-cycle (and learned-cycle) parameters: the coarsest level is solved to
    1e-2 relative residual per cycle; solving it further changes neither the
    cycle count nor the final residual at J5..J7 (1e-10: 11 cycles, 15.5 ms;
    1e-2: 11 cycles, 5.0 ms), only the cost."""
    return MgParams.defaults(coarse_tol=1e-2, fine_tol=tol, coarse_iter=400, max_vcycle=max_vcycle)


def harvest(prob: Problem, ncycles: int = 6):
    """training rows (X [n, 192], T [n, 64]) of one draw"""
    top = prob.top
    p = MgParams.defaults(coarse_tol=1e-12, fine_tol=1e-13, coarse_iter=400, max_vcycle=60)
    prob.reset()
    prob.h.solve(SOLVE_FMG, p)
    ustar = top.fs_download(FS_ETA)
    f = top.fs_download(FS_RHS)
    prob.reset()
    one = vcycle_params(tol=0.0, max_vcycle=1)
    X, T = [], []
    for _ in range(ncycles):
        top.elliptic_apply(FS_ETA, FS_DIVF)
        r = f - top.fs_download(FS_DIVF)
        top.fs_upload(FS_DIVF_OLD, r)
        feats, sigma, cells = prob.h.features(FS_DIVF_OLD)
        u = top.fs_download(FS_ETA)
        keep = sigma > 0
        X.append(feats[keep])
        T.append(((ustar - u)[cells] / sigma[:, None])[keep].astype(np.float32))
        prob.h.solve(SOLVE_VCYCLE, one)  # advances u by one V-cycle
    return np.concatenate(X), np.concatenate(T)


def compare(prob: Problem, mlp: MLP, tol=1e-9, max_vcycle=30):
    """(plain V-cycles, learned) -> (cycles, final residual, ms) each"""
    p = vcycle_params(tol, max_vcycle)
    prob.h.set_learned(None)
    prob.solve(SOLVE_VCYCLE, p)  # warm
    it0, r0, ms0 = prob.solve(SOLVE_VCYCLE, p)
    prob.h.set_learned(mlp)
    prob.solve(SOLVE_LEARNED, p)
    it1, r1, ms1 = prob.solve(SOLVE_LEARNED, p)
    prob.h.set_learned(None)
    return (int(it0[-1]), float(r0[-1]), ms0), (int(it1[-1]), float(r1[-1]), ms1)


def train_network(X, T, steps=3000, batch=1024, lr=1e-3, seed=0, max_batch=None):
    mlp = MLP(max_batch=max(batch, max_batch or 0), seed=seed)
    mlp.set_dataset(X, T)
    losses = mlp.train(steps, batch, lr=lr, seed=seed + 1)
    return mlp, losses


def load_network(max_batch: int, path: str = WEIGHTS):
    if not os.path.exists(path):
        return None
    mlp = MLP(max_batch=max_batch, seed=0)
    mlp.params = np.load(path)["params"]
    return mlp


def main(ndraws: int = 50, npairs: int = 200_000, steps: int = 4000, fine: int = 7, out: str = WEIGHTS):
    grids = hierarchy_grids(fine)
    X, T = [], []
    t0 = time.perf_counter()
    for s in range(ndraws):
        prob = Problem(grids, seed=1000 + s)
        x, t = harvest(prob)
        prob.close()
        X.append(x)
        T.append(t)
        print(f"draw {s}: {len(x)} pairs ({time.perf_counter() - t0:.1f} s)", flush=True)
    X, T = np.concatenate(X), np.concatenate(T)
    sel = np.random.default_rng(0).permutation(len(X))[:npairs]
    X, T = X[sel], T[sel]
    npatch = 10 * (grids[-1].layout.n // 8) ** 2
    mlp, losses = train_network(X, T, steps=steps, max_batch=npatch)
    print(f"trained on {len(X)} pairs: loss {losses[:10].mean():.4f} -> {losses[-50:].mean():.4f}", flush=True)
    np.savez(out, params=mlp.params, losses=losses, ndraws=ndraws, npairs=len(X), steps=steps, fine=fine)
    for s in range(3):  # held-out draws
        prob = Problem(grids, seed=5000 + s)
        (i0, r0, m0), (i1, r1, m1) = compare(prob, mlp)
        prob.close()
        print(f"held-out {s}: V-cycles {i0} ({r0:.2e}, {m0:.1f} ms)  learned {i1} ({r1:.2e}, {m1:.1f} ms)", flush=True)
    return out


if __name__ == "__main__":
    import sys
    main(*(int(a) for a in sys.argv[1:4]))
