"""FP32-bulk split-explicit steps (RK4_split with the trend stencils in FP32)
against the FP64 oracle (needs a GPU).

In the FP32-bulk mode the four trend evaluations of RK4_split run the FP32
stencils of the FP32-bulk stage kernel and write FP32 trends; the stage
states, scalar_star/u_star, the barotropic correction, the flux divergence
and the implicit free surface stay FP64.  Bars, as max-norm errors against
the oracle's FP64 split steps on the same inputs:

* 1 step (J5 ocean, 4 layers, dt 1800 s, theta1 0.6, theta2 0.8, BiCGSTAB to
  1e-12): free surface within 1e-8 of the column depth, mass within 1e-8 and
  velocity within 1e-6 relative;
* 100 steps: free surface within 1e-6 of the column depth, mass within 1e-6,
  velocity within 1e-5 relative;
* the same step on a J4..J6 hierarchy with V-cycles to 1e-12: within 1e-7
  (free surface, mass) and 1e-5 (velocity) of the oracle's single-level step.
A dropped or mis-signed term in an FP32 stencil moves these errors to 1e-3
or more.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from oracle.oracle import EllipticOracle, geometry_of

pytestmark = pytest.mark.gpu

RHO0 = 1028.0
DT, TH1, TH2 = 1800.0, 0.6, 0.8


def _errors(g, got, want, topo):
    m, _, u, e = got
    M, _, U, E = want
    own = g.layout.owned_cells()
    own_e = g.layout.owned_cells(with_poles=False)
    depth = np.max(M[:, own].sum(0) / RHO0)
    return {"mass": np.max(np.abs(m[:, own] - M[:, own])) / np.max(np.abs(M[:, own])),
            "velo": np.max(np.abs(u[:, :, own_e] - U[:, :, own_e])) / np.max(np.abs(U[:, :, own_e])),
            "eta/depth": np.max(np.abs(e[own] - E[own])) / depth}


def _start(g, L, seed):
    topo, mass, temp, velo = cases.ocean(g, L, seed=seed)
    return topo, mass, temp, velo, mass.sum(0) / RHO0 + topo


def test_fp32_bulk_split_against_oracle():
    print("FP32-bulk split steps vs oracle: J5 L4", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L = 4
    topo, mass, temp, velo, eta = _start(g, L, 31)
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=DT, theta1=TH1, theta2=TH2)
    ms, ts, us = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(ms, ts, us)
    want = (ms, ts, us, o.node_ghost(eta))
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.fs_set_params(DT, TH1, TH2)
    ctx.upload_state(0, mass, temp, velo)
    ctx.fs_upload(W.FS_ETA, eta)
    ctx.set_precision(W.PRECISION_FP32_BULK)
    done = 0
    for nsteps, bars in ((1, {"mass": 1e-8, "velo": 1e-6, "eta/depth": 1e-8}),
                         (100, {"mass": 1e-6, "velo": 1e-5, "eta/depth": 1e-6})):
        while done < nsteps:
            *want, _ = o.rk4_split(*want, tol=1e-12, maxit=400)
            ctx.rk4_split(0, 1, 1e-12, 400)
            done += 1
        got = (*ctx.download_state(0), ctx.fs_download(W.FS_ETA))
        err = _errors(g, got, want, topo)
        print(f"  {nsteps} steps: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
        for k, v in err.items():
            assert v <= bars[k], f"{nsteps} steps: {k} error {v:.3e} above {bars[k]:.0e}"
        assert err["velo"] > 0  # the FP32 stencils are in effect
    assert ctx.stage_shape()[0] != W.SHAPE_AUTO  # tuned on the first FP32-bulk step


def test_fp32_bulk_split_on_hierarchy():
    print("FP32-bulk split step on a J4..J6 hierarchy (V-cycles)", flush=True)
    grids = [W.Grid(level=l, min_level=4, domain_level=1) for l in (4, 5, 6)]
    L = 3
    ctxs = []
    for k, g in enumerate(grids):
        topo, mass, temp, velo, eta = _start(g, L, 21)
        c = W.Context(g, nlayers=L if k == 2 else 1)
        c.set_topography(topo)
        c.fs_set_params(DT, TH1, TH2)
        ctxs.append(c)
    g, top = grids[-1], ctxs[-1]
    top.upload_state(0, mass, temp, velo)
    top.fs_upload(W.FS_ETA, eta)
    top.set_precision(W.PRECISION_FP32_BULK)
    h = W.Hierarchy(ctxs)
    try:
        p = W.MgParams.defaults(coarse_tol=1e-12, fine_tol=1e-12, coarse_iter=400, max_vcycle=60)
        it = h.rk4_split(0, 1, W.SOLVE_VCYCLE, p)
    finally:
        h.close()
    assert 0 < it < 60
    got = (*top.download_state(0), top.fs_download(W.FS_ETA))
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=DT, theta1=TH1, theta2=TH2)
    ms, ts, us = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(ms, ts, us)
    *want, _ = o.rk4_split(ms, ts, us, o.node_ghost(eta), tol=1e-13, maxit=400)
    err = _errors(g, got, want, topo)
    print(f"  {it} cycles: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["eta/depth"] <= 1e-7 and err["mass"] <= 1e-7 and err["velo"] <= 1e-5


def test_fp32_bulk_split_J8_L30_against_fp64():
    print("FP32-bulk vs FP64 split steps: J8 L30, 10 steps", flush=True)
    g = W.Grid(level=8, min_level=8, domain_level=2)
    L = 30
    topo, mass, temp, velo, eta = _start(g, L, 7)
    res = {}
    for prec in (W.PRECISION_FP64, W.PRECISION_FP32_BULK):
        ctx = W.Context(g, nlayers=L)
        ctx.set_topography(topo)
        ctx.fs_set_params(900.0, TH1, TH2)
        ctx.upload_state(0, mass, temp, velo)
        ctx.fs_upload(W.FS_ETA, eta)
        ctx.set_precision(prec)
        ctx.rk4_split(0, 10, 1e-9, 400)
        res[prec] = (*ctx.download_state(0), ctx.fs_download(W.FS_ETA))
        ctx.close()
    err = _errors(g, res[W.PRECISION_FP32_BULK], res[W.PRECISION_FP64], topo)
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["eta/depth"] <= 1e-6 and err["mass"] <= 1e-6 and err["velo"] <= 1e-5


def test_config2_j7_l20_fp32_bulk_100_steps_against_fp64():
    """Config 2 at its workload: the J7 ocean with 20 layers, split steps with
    the free surface solved by V-cycles on the J5..J7 hierarchy to 1e-9
    relative residual every step (dt 300 s, theta1 0.6, theta2 0.8), 100
    steps; FP64 (stencils, smoothing) against FP32-bulk stencils with FP32
    smoothing and FP64 residual refinement.  The FP64 path is the oracle's
    step to 1e-11 on this hierarchy (tests/test_gpu_ml.py), so its 100 steps
    stand for the oracle's here (a C run of 100 J7 L20 steps takes minutes).
    Every step converges (fewer cycles than the cap); after 100 steps the
    free surface is within 1e-6 of the column depth, mass within 1e-6 and
    velocity within 1e-5 of the FP64 run."""
    print("config 2: J5..J7 L20, 100 split steps, V-cycles to 1e-9, FP32-bulk vs FP64", flush=True)
    grids = [W.Grid(level=l, min_level=5, domain_level=2) for l in (5, 6, 7)]
    L, dt, nsteps, cap = 20, 300.0, 100, 40
    draws = [_start(g, L if k == 2 else 1, 17) for k, g in enumerate(grids)]
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-9, coarse_iter=400, max_vcycle=cap)
    res, cycles = {}, {}
    for prec in (W.PRECISION_FP64, W.PRECISION_FP32_BULK):
        ctxs = []
        for k, g in enumerate(grids):
            c = W.Context(g, nlayers=L if k == 2 else 1)
            c.set_topography(draws[k][0])
            c.fs_set_params(dt, TH1, TH2)
            ctxs.append(c)
        top = ctxs[-1]
        topo, mass, temp, velo, eta = draws[2]
        top.upload_state(0, mass, temp, velo)
        top.fs_upload(W.FS_ETA, eta)
        top.set_precision(prec)
        h = W.Hierarchy(ctxs)
        h.set_precision(prec)
        try:
            cycles[prec] = [h.rk4_split(0, 1, W.SOLVE_VCYCLE, p) for _ in range(nsteps)]
        finally:
            h.close()
        res[prec] = (*top.download_state(0), top.fs_download(W.FS_ETA))
        for c in ctxs:
            c.close()
    for prec, cyc in cycles.items():
        assert all(0 < n < cap for n in cyc), (prec, cyc)
    err = _errors(grids[-1], res[W.PRECISION_FP32_BULK], res[W.PRECISION_FP64], draws[2][0])
    print(f"  cycles per step FP64 {min(cycles[0])}-{max(cycles[0])}, FP32-bulk {min(cycles[1])}-{max(cycles[1])}; "
          + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["eta/depth"] <= 1e-6 and err["mass"] <= 1e-6 and err["velo"] <= 1e-5
