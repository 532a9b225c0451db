"""Laplacian diffusion and hyperdiffusion on the GPU (needs a GPU).

wtg_set_diffusion switches the trend's diffusion hooks on; the pre-pass
kernels (div u, Laplacians, vorticity, rot rot) and the stage kernel's
DIFF instantiation against the oracle and the reference.

Bars (relative to the field's max-norm over the owned cells, per layer):
  trend vs oracle          1e-14 (same evaluation order, no FMA contraction;
                           observed bit-identical)
  trend, RK4 vs reference  1e-13 (fixtures trisk_ocean_diff{1,2}_J4)
  RK4 steps vs oracle      1e-13, synthetic J5 ocean, orders in every mix
  graph replay, ensembles  bit-identical to eager / single-member runs
  split step vs oracle     1e-12 with the same BiCGSTAB iteration count (the
                           GPU's norms reduce in another order)
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from conftest import golden, maxrel
from helpers import fixture_state, grid_of_fixture, oracle_of

pytestmark = pytest.mark.gpu

FIXTURES = ["ocean_diff1_J4", "ocean_diff2_J4"]


def _cmp(got, want, layout, tol, what):
    own, own_e = layout.owned_cells(), layout.owned_edge_cells()
    m, t, u = got
    M, T, U = want
    for k in range(m.shape[0]):
        assert maxrel(m[k, own], M[k, own]) <= tol, f"{what} mass layer {k}"
        assert maxrel(t[k, own], T[k, own]) <= tol, f"{what} temp layer {k}"
        for e in range(3):
            assert maxrel(u[k, e, own_e], U[k, e, own_e]) <= tol, f"{what} velo layer {k} edge {e}"


def _fixture_ctx(f, g, topo, mass, temp, velo):
    L = int(f["nlayers"])
    ctx = W.Context(g, nlayers=L, grav=float(f["grav"]), ref_density=float(f["ref_density"]))
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    lp = [int(x) for x in f["laplace"]]
    vs = [float(x) for x in f["visc"]]
    ctx.set_diffusion(lp, vs)
    return ctx, (*lp, *vs)


@pytest.mark.parametrize("case", FIXTURES)
def test_trend_with_diffusion(case):
    print(f"diffusion trend {case}", flush=True)
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    ctx, diff = _fixture_ctx(f, g, topo, mass, temp, velo)
    ctx.trend(0, 1)
    got = ctx.download_state(1)
    o = oracle_of(g, topo, f)
    m, t, u = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(m, t, u)
    _cmp(got, o.trend(m, t, u, diffusion=diff), g.layout, 1e-14, "trend vs oracle")
    ref = tuple(np.nan_to_num(f[k]) for k in ("dmass", "dtemp", "dvelo"))
    _cmp(got, ref, g.layout, 1e-13, "trend vs reference")
    ctx.close()


@pytest.mark.parametrize("case", FIXTURES)
def test_rk4_with_diffusion_matches_reference(case):
    print(f"diffusion RK4 {case}", flush=True)
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    ctx, _ = _fixture_ctx(f, g, topo, mass, temp, velo)
    for s in range(1, int(f["nsteps"]) + 1):
        ctx.rk4(0, float(f["dt"]), 1)
        ref = tuple(np.nan_to_num(f[f"{v}{s}"]) for v in ("mass", "temp", "velo"))
        _cmp(ctx.download_state(0), ref, g.layout, 1e-13, f"step {s} vs reference")
    ctx.close()


# J5 (dx ~ 220 km), dt 600 s: viscosities ~1e-2 dx^(2 Lap) / dt
VISC = {1: (8e5, 4e5, 1.2e6, 8e5), 2: (4e16, 2e16, 6e16, 4e16)}
MIXES = [(1, 1, 1), (2, 2, 2), (2, 1, 2), (1, 0, 2), (0, 2, 1), (2, 0, 0)]


def _visc(mix):
    ls, ld, lr = mix
    return (VISC[ls][0] if ls else 0.0, VISC[ls][1] if ls else 0.0, VISC[ld][2] if ld else 0.0,
            VISC[lr][3] if lr else 0.0)


@pytest.mark.parametrize("mix", MIXES)
def test_rk4_with_diffusion_matches_oracle(mix):
    print(f"diffusion RK4 J5 L3 orders {mix}", flush=True)
    from oracle.oracle import Oracle, geometry_of
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, dt = 3, 600.0
    topo, mass, temp, velo = cases.ocean(g, L, seed=11)
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    ctx.set_diffusion(mix, _visc(mix))
    ctx.rk4(0, dt, 3)
    got = ctx.download_state(0)
    o = Oracle(g.layout, geometry_of(g), topo)
    want = o.rk4(mass, temp, velo, dt, 3, diffusion=(*mix, *_visc(mix)))
    _cmp(got, want, g.layout, 1e-13, "3 steps vs oracle")
    # the diffusion changes the step (the mix is not a no-op)
    plain = o.rk4(mass, temp, velo, dt, 3)
    own_e = g.layout.owned_edge_cells()
    assert maxrel(got[2][:, :, own_e], plain[2][:, :, own_e]) > 1e-9 or maxrel(got[0], plain[0]) > 1e-9
    ctx.close()


def test_diffusion_graph_replay_equals_eager():
    print("diffusion RK4 graph replay J5 L3", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, dt, mix = 3, 600.0, (2, 2, 2)
    topo, mass, temp, velo = cases.ocean(g, L, seed=12)
    out = []
    for graph in (False, True):
        ctx = W.Context(g, nlayers=L)
        ctx.set_topography(topo)
        ctx.upload_state(0, mass, temp, velo)
        ctx.set_diffusion(mix, _visc(mix))
        ctx.set_graph(graph)
        ctx.rk4(0, dt, 2)
        ctx.rk4(0, dt, 2)
        out.append(ctx.download_state(0))
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_diffusion_ensemble_members_equal_single_runs():
    print("diffusion RK4 ensemble of 2, J5 L3", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, dt, mix = 3, 600.0, (2, 2, 2)
    draws = [cases.ocean(g, L, seed=20 + k) for k in range(2)]
    ens = W.Context(g, nlayers=L, members=2)
    ens.set_topography(np.stack([d[0] for d in draws]))
    ens.upload_state(0, *(np.concatenate([d[v] for d in draws]) for v in (1, 2, 3)))
    ens.set_diffusion(mix, _visc(mix))
    ens.rk4(0, dt, 2)
    em, et, eu = ens.download_state(0)
    ens.close()
    for k, (topo, m, t, u) in enumerate(draws):
        c = W.Context(g, nlayers=L)
        c.set_topography(topo)
        c.upload_state(0, m, t, u)
        c.set_diffusion(mix, _visc(mix))
        c.rk4(0, dt, 2)
        sm, st, su = c.download_state(0)
        c.close()
        sl = slice(k * L, (k + 1) * L)
        assert np.array_equal(em[sl], sm) and np.array_equal(et[sl], st) and np.array_equal(eu[sl], su)


@pytest.mark.parametrize("fused", [True, False])
def test_split_step_with_diffusion_matches_oracle(fused, monkeypatch):
    print(f"diffusion RK4_split J4 L3 (fused stage {fused})", flush=True)
    from oracle.oracle import EllipticOracle, geometry_of
    monkeypatch.setenv("WTG_SPLIT_FUSED", "1" if fused else "0")
    g = W.Grid(level=4, min_level=4, domain_level=2)
    L, dt, th1, th2, mix = 3, 1200.0, 0.6, 0.8, (2, 2, 2)
    visc = (4 * v for v in _visc(mix))  # dx doubles from J5 to J4: 16x for order 2, kept moderate
    visc = tuple(visc)
    topo, mass, temp, velo = cases.ocean(g, L, seed=13)
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=dt, theta1=th1, theta2=th2)
    m, t, u = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(m, t, u)
    eta = o.node_ghost(m.sum(0) / 1028.0 + topo)
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.fs_set_params(dt, th1, th2)
    ctx.upload_state(0, m, t, u)
    ctx.fs_upload(W.FS_ETA, eta)
    ctx.set_diffusion(mix, visc)
    it = ctx.rk4_split(0, 1, 1e-13, 400)
    m2, t2, u2 = ctx.download_state(0)
    e2 = ctx.fs_download(W.FS_ETA)
    ctx.close()
    mo, to, uo, eo, ito = o.rk4_split(m, t, u, eta, tol=1e-13, maxit=400, diffusion=(*mix, *visc))
    print(f"  BiCGSTAB iterations: GPU {it}, oracle {ito}", flush=True)
    assert it == ito
    masters = g.layout.owned_cells()
    own = g.layout.owned_cells(with_poles=False)
    for a, b, cells in ((m2, mo, masters), (t2, to, masters), (u2, uo, own)):
        assert maxrel(a[..., cells], b[..., cells]) <= 1e-12
    assert maxrel(e2[masters], eo[masters]) <= 1e-12


def test_diffusion_refusals():
    print("diffusion refusals", flush=True)
    g = W.Grid(level=4, min_level=4, domain_level=2)
    topo, mass, temp, velo = cases.ocean(g, 2, seed=1)
    ctx = W.Context(g, nlayers=2)
    ctx.set_topography(topo)
    with pytest.raises(W.WtgError):
        ctx.set_diffusion((1, 1, 0), (1.0, 1.0, 1.0, 0.0))  # div u term without the rot u order
    with pytest.raises(W.WtgError):
        ctx.set_diffusion((3, 0, 0), (1.0, 1.0, 0.0, 0.0))
    ctx.set_diffusion((2, 2, 2), (1.0, 1.0, 1.0, 1.0))
    with pytest.raises(W.WtgError):
        ctx.set_precision(W.PRECISION_FP32_BULK)
    ctx.set_diffusion((0, 0, 0))
    ctx.set_precision(W.PRECISION_FP32_BULK)
    with pytest.raises(W.WtgError):
        ctx.set_diffusion((1, 0, 0), (1.0, 1.0, 0.0, 0.0))
    ctx.close()
    grids = [W.Grid(level=l, min_level=4, domain_level=2) for l in (4, 5)]
    ctxs = []
    for gg in grids:
        c = W.Context(gg, nlayers=1)
        c.set_topography(np.zeros(gg.layout.ncells) - 4000.0)
        ctxs.append(c)
    ctxs[1].set_diffusion((1, 0, 1), (1.0, 1.0, 0.0, 1.0))
    ml = W.MultiLevel(ctxs)
    try:
        with pytest.raises(W.WtgError):
            ml.rk4(0, 600.0, 1)
    finally:
        ml.close()
        for c in ctxs:
            c.close()
