"""HIP split-explicit step with implicit free surface against the CPU oracle
and the reference (needs a GPU).

Bars: one RK4_split step (J4, 3 layers, dt 1200 s, theta1 0.6, theta2 0.8,
BiCGSTAB to 1e-13) within 1e-12 of the reference (tests/golden/
elliptic_splitJ4.npz) and of the oracle in every field -- the GPU's norms
reduce in a different order, so the two solvers' iterates differ in the last
bits; with the same BiCGSTAB iteration count.  A J8 ocean step: columns
dilated to the new free surface and the free surface's area integral
conserved, as the oracle's properties state.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden, maxrel
from helpers import grid_of_fixture
from oracle.oracle import EllipticOracle, geometry_of

pytestmark = pytest.mark.gpu


def test_split_step_matches_reference_and_oracle():
    print("split step J4", flush=True)
    f = golden("elliptic_splitJ4.npz")
    g = grid_of_fixture(f)
    topo = np.nan_to_num(f["topo"])
    kw = dict(grav=float(f["grav"]), rho0=float(f["ref_density"]), dt=float(f["dt"]), theta1=float(f["theta1"]),
              theta2=float(f["theta2"]))
    o = EllipticOracle(g.layout, geometry_of(g), topo, **kw)
    mass, temp, velo = (np.nan_to_num(f[k]).copy() for k in ("mass", "temp0", "velo"))
    o.ghost_fill(mass, temp, velo)
    eta = o.node_ghost(np.nan_to_num(f["eta_old"]))
    L = mass.shape[0]
    ctx = W.Context(g, nlayers=L, grav=kw["grav"], ref_density=kw["rho0"])
    ctx.set_topography(topo)
    ctx.fs_set_params(kw["dt"], kw["theta1"], kw["theta2"])
    ctx.upload_state(0, mass, temp, velo)
    ctx.fs_upload(W.FS_ETA, eta)
    it = ctx.rk4_split(0, 1, float(f["tol"]), 400)
    m2, t2, u2 = ctx.download_state(0)
    e2 = ctx.fs_download(W.FS_ETA)
    mo, to, uo, eo, ito = o.rk4_split(mass, temp, velo, eta, tol=float(f["tol"]), maxit=400)
    print(f"BiCGSTAB iterations: GPU {it}, oracle {ito}", flush=True)
    assert it == ito
    masters = g.layout.owned_cells()
    own = g.layout.owned_cells(with_poles=False)
    for a, b, r, cells in ((m2, mo, f["sp_mass"], masters), (t2, to, f["sp_temp"], masters),
                           (u2, uo, f["sp_velo"], own)):
        assert maxrel(a[..., cells], r[..., cells]) <= 1e-12
        assert maxrel(a[..., cells], b[..., cells]) <= 1e-12
    assert maxrel(e2[masters], f["sp_eta"][masters]) <= 1e-12
    assert maxrel(e2[masters], eo[masters]) <= 1e-12


def test_split_steps_J8_properties():
    print("split steps J8 L30", flush=True)
    from wavetrisk_hydrostatic_amd import cases
    g = W.Grid(level=8, min_level=8, domain_level=2)
    L = 30
    topo, mass, temp, velo = cases.ocean(g, L, seed=7)
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.fs_set_params(900.0, 0.6, 0.8)
    ctx.upload_state(0, mass, temp, velo)
    eta0 = mass.sum(0) / 1028.0 + topo  # the state's own free surface
    ctx.fs_upload(W.FS_ETA, eta0)
    it = ctx.rk4_split(0, 3, 1e-11, 400)
    m2, _, u2 = ctx.download_state(0)
    e2 = ctx.fs_download(W.FS_ETA)
    own = g.layout.owned_cells()  # owned nodes and the two pole masters, each node once
    assert np.isfinite(m2[:, own]).all() and np.isfinite(u2[:, :, own]).all() and it <= 400
    thick = m2[:, own].sum(0) / 1028.0
    assert np.max(np.abs(thick - (e2[own] - topo[own]))) <= 1e-11 * 4000.0
    area = 1.0 / g.array("hex_inv")[own]
    d = abs(np.sum(area * (e2[own] - eta0[own])))
    print(f"J8: BiCGSTAB {it} iterations, relative change of the eta integral {d / np.sum(area * np.abs(eta0[own])):.2e}",
          flush=True)
    assert d <= 1e-8 * np.sum(area * np.abs(eta0[own])), d


@pytest.mark.parametrize("method", ["vcycle", "learned"])
def test_split_step_on_hierarchy(method):
    """The same step with the free surface solved on a J4..J6 hierarchy by
    V-cycles (or learned cycles, any network) to 1e-12 agrees with the oracle's
    single-level BiCGSTAB step to 1e-13 within 1e-9."""
    print(f"split step on a hierarchy ({method})", flush=True)
    from wavetrisk_hydrostatic_amd import cases
    grids = [W.Grid(level=l, min_level=4, domain_level=1) for l in (4, 5, 6)]
    L, dt = 3, 1800.0
    ctxs = []
    for k, g in enumerate(grids):
        topo, mass, temp, velo = cases.ocean(g, L, seed=21)
        c = W.Context(g, nlayers=L if k == 2 else 1)
        c.set_topography(topo)
        c.fs_set_params(dt, 0.6, 0.8)
        ctxs.append(c)
    g, top = grids[-1], ctxs[-1]
    eta = mass.sum(0) / 1028.0 + topo
    top.upload_state(0, mass, temp, velo)
    top.fs_upload(W.FS_ETA, eta)
    h = W.Hierarchy(ctxs)
    mlp = None
    try:
        if method == "learned":
            mlp = W.MLP(max_batch=h.npatches, seed=5)
            h.set_learned(mlp)
        p = W.MgParams.defaults(coarse_tol=1e-12, fine_tol=1e-12, coarse_iter=400, max_vcycle=60)
        it = h.rk4_split(0, 1, W.SOLVE_VCYCLE if method == "vcycle" else W.SOLVE_LEARNED, p)
    finally:
        h.close()
        if mlp is not None:
            mlp.close()
    print(f"{method}: {it} cycles", flush=True)
    assert 0 < it < 60
    m2, t2, u2 = top.download_state(0)
    e2 = top.fs_download(W.FS_ETA)
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=dt, theta1=0.6, theta2=0.8)
    mg_, tg_, ug_ = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(mg_, tg_, ug_)
    mo, to, uo, eo, _ = o.rk4_split(mg_, tg_, ug_, o.node_ghost(eta), tol=1e-13, maxit=400)
    masters = g.layout.owned_cells()
    own = g.layout.owned_cells(with_poles=False)
    assert maxrel(e2[masters], eo[masters]) <= 1e-9
    assert maxrel(m2[:, masters], mo[:, masters]) <= 1e-9
    assert maxrel(u2[:, :, own], uo[:, :, own]) <= 1e-9


def test_fused_split_stage_equals_trend_plus_axpy():
    """The FP64 split-explicit stage fused into the stage kernel's update
    gives the same bits as the trend launch plus split_axpy_kernel (contexts
    created with WTG_SPLIT_FUSED=0 take the unfused path): J5 L4, two steps,
    a single context and a two-member ensemble."""
    import os
    print("fused split stage vs trend + AXPY, J5 L4", flush=True)
    from wavetrisk_hydrostatic_amd import cases
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L = 4
    topo, mass, temp, velo = cases.ocean(g, L, seed=13)
    eta = mass.sum(0) / 1028.0 + topo
    out = {}
    for fused in (True, False):
        if fused:
            os.environ.pop("WTG_SPLIT_FUSED", None)
        else:
            os.environ["WTG_SPLIT_FUSED"] = "0"
        try:
            ctx = W.Context(g, nlayers=L)
            ens = W.Context(g, nlayers=L, members=2)
        finally:
            os.environ.pop("WTG_SPLIT_FUSED", None)
        ctx.set_topography(topo)
        ctx.fs_set_params(900.0, 0.6, 0.8)
        ctx.upload_state(0, mass, temp, velo)
        ctx.fs_upload(W.FS_ETA, eta)
        it = ctx.rk4_split(0, 2, 1e-11, 400)
        ens.set_topography(np.stack([topo, topo]))
        ens.fs_set_params(900.0, 0.6, 0.8)
        ens.upload_state(0, np.concatenate([mass, mass]), np.concatenate([temp, temp]),
                         np.concatenate([velo, velo]))
        ens.fs_upload(W.FS_ETA, np.stack([eta, eta]))
        ite = ens.rk4_split(0, 2, 1e-11, 400)
        out[fused] = (it, ite, ctx.download_state(0), ctx.fs_download(W.FS_ETA), ens.download_state(0),
                      ens.fs_download(W.FS_ETA))
        ctx.close()
        ens.close()
    a, b = out[True], out[False]
    assert a[0] == b[0] and a[1] == b[1]
    own, owne = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for x, y in ((a[2], b[2]), (a[4], b[4])):
        assert np.array_equal(x[0][..., own], y[0][..., own]) and np.array_equal(x[1][..., own], y[1][..., own])
        assert np.array_equal(x[2][..., owne], y[2][..., owne])
    assert np.array_equal(a[3][own], b[3][own]) and np.array_equal(a[5][:, own], b[5][:, own])


@pytest.mark.parametrize("nstages", [3, 2, 1])
def test_other_split_schemes_match_reference_and_oracle(nstages):
    """RK3_split, RK2_split, Euler_split (wtg_rk_split) on the J4 case of
    tests/golden/split_schemes_J4.npz: the oracle's BiCGSTAB count, the
    state within 1e-12 and the free surface within 5e-12 of the reference
    and of the oracle (the bars of tests/test_split_oracle.py)"""
    print(f"split step J4, {nstages} stages", flush=True)
    f = golden("elliptic_splitJ4.npz")
    s = golden("split_schemes_J4.npz")
    g = grid_of_fixture(f)
    topo = np.nan_to_num(f["topo"])
    kw = dict(grav=float(f["grav"]), rho0=float(f["ref_density"]), dt=float(f["dt"]), theta1=float(f["theta1"]),
              theta2=float(f["theta2"]))
    o = EllipticOracle(g.layout, geometry_of(g), topo, **kw)
    mass, temp, velo = (np.nan_to_num(f[k]).copy() for k in ("mass", "temp0", "velo"))
    o.ghost_fill(mass, temp, velo)
    eta = o.node_ghost(np.nan_to_num(f["eta_old"]))
    ctx = W.Context(g, nlayers=mass.shape[0], grav=kw["grav"], ref_density=kw["rho0"])
    ctx.set_topography(topo)
    ctx.fs_set_params(kw["dt"], kw["theta1"], kw["theta2"])
    ctx.upload_state(0, mass, temp, velo)
    ctx.fs_upload(W.FS_ETA, eta)
    it = ctx.rk4_split(0, 1, float(f["tol"]), 400, nstages=nstages)
    m2, t2, u2 = ctx.download_state(0)
    e2 = ctx.fs_download(W.FS_ETA)
    mo, to, uo, eo, ito = o.rk4_split(mass, temp, velo, eta, tol=float(f["tol"]), maxit=400, nstages=nstages)
    assert it == ito
    masters = g.layout.owned_cells()
    own = g.layout.owned_cells(with_poles=False)
    p = f"sp{nstages}_"
    for a, b, r, cells in ((m2, mo, s[p + "mass"], masters), (t2, to, s[p + "temp"], masters),
                           (u2, uo, s[p + "velo"], own)):
        assert maxrel(a[..., cells], r[..., cells]) <= 1e-12
        assert maxrel(a[..., cells], b[..., cells]) <= 1e-12
    assert maxrel(e2[masters], s[p + "eta"][masters]) <= 5e-12
    assert maxrel(e2[masters], eo[masters]) <= 5e-12
    ctx.close()
