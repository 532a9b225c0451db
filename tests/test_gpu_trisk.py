"""HIP TRiSK path against the CPU oracle and the reference (needs a GPU).

Bar: FP64 fields within 1e-12 relative of the reference (max-norm over a
field); against the oracle on identical inputs the HIP kernels are expected
to agree to the last bit (same evaluation order, no FMA contraction), and
the assertion allows 1e-14.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from conftest import golden, maxrel
from helpers import fixture_state, grid_of_fixture, oracle_of

pytestmark = pytest.mark.gpu

CASES = ["tc5_J4", "ocean_J4", "ident_J4", "tc5_J5"]
TOL_REF = 1e-12
TOL_ORACLE = 1e-14


def _ctx(grid, f, nlayers):
    return W.Context(grid, nlayers=nlayers, grav=float(f["grav"]), ref_density=float(f["ref_density"]))


def _cmp_state(got, want, layout, tol, what):
    own, own_e = layout.owned_cells(), layout.owned_edge_cells()
    m, t, u = got
    M, T, U = want
    for k in range(m.shape[0]):
        assert maxrel(m[k, own], M[k, own]) <= tol, f"{what} mass layer {k}"
        if np.any(T[k, own]):
            assert maxrel(t[k, own], T[k, own]) <= tol, f"{what} temp layer {k}"
        for e in range(3):
            assert maxrel(u[k, e, own_e], U[k, e, own_e]) <= tol, f"{what} velo layer {k} edge {e}"


@pytest.mark.parametrize("case", CASES)
def test_trend_matches_oracle_and_reference(case):
    f = golden(f"trisk_{case}.npz")
    print(f"case {case}", flush=True)
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    L = int(f["nlayers"])
    ctx = _ctx(g, f, L)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    ctx.trend(0, 1)
    got = ctx.download_state(1)
    o = oracle_of(g, topo, f)
    m, t, u = mass.copy(), temp.copy(), velo.copy()
    o.ghost_fill(m, t, u)
    want = o.trend(m, t, u)
    _cmp_state(got, want, g.layout, TOL_ORACLE, "trend vs oracle")
    ref = (np.nan_to_num(f["dmass"]), np.nan_to_num(f["dtemp"]), np.nan_to_num(f["dvelo"]))
    _cmp_state(got, ref, g.layout, TOL_REF, "trend vs reference")


@pytest.mark.parametrize("case", CASES)
def test_rk4_matches_reference(case):
    f = golden(f"trisk_{case}.npz")
    print(f"case {case}", flush=True)
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    L = int(f["nlayers"])
    ctx = _ctx(g, f, L)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    for s in range(1, int(f["nsteps"]) + 1):
        ctx.rk4(0, float(f["dt"]), 1)
        got = ctx.download_state(0)
        ref = tuple(np.nan_to_num(f[f"{v}{s}"]) for v in ("mass", "temp", "velo"))
        _cmp_state(got, ref, g.layout, TOL_REF, f"step {s} vs reference")


def test_c1_100_steps_match_reference_and_oracle():
    """C1 at its workload: J5 TC5, 100 RK4 steps at CFL 0.8 from the
    reference's state; within 1e-12 of the reference after steps 25, 50, 75,
    100 and within 1e-13 of the oracle's 100 steps"""
    f = golden("trisk_tc5_c1_J5.npz")
    print("C1: J5 TC5 100 RK4 steps", flush=True)
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    ctx = _ctx(g, f, 1)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    every, dt = int(f["every"]), float(f["dt"])
    for s in range(1, int(f["nsteps"]) + 1):
        ctx.rk4(0, dt, every)
        got = ctx.download_state(0)
        ref = (f[f"mass{s}"], np.nan_to_num(f[f"temp{s}"]), f[f"velo{s}"])
        _cmp_state(got, ref, g.layout, TOL_REF, f"step {s * every} vs reference")
    o = oracle_of(g, topo, f)
    want = o.rk4(mass, temp, velo, dt, every * int(f["nsteps"]))
    _cmp_state(got, want, g.layout, 1e-13, "100 steps vs oracle")
    ctx.close()


@pytest.mark.parametrize("level,dl,nlayers,steps", [(6, 3, 5, 3), (7, 2, 2, 2)])
def test_rk4_matches_oracle_synthetic(level, dl, nlayers, steps):
    print(f"J{level} L{nlayers}", flush=True)
    g = W.Grid(level=level, min_level=level, domain_level=dl)
    topo, m, t, u = cases.ocean(g, nlayers, seed=level)
    ctx = W.Context(g, nlayers=nlayers)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.rk4(0, 600.0, steps)
    got = ctx.download_state(0)
    o = oracle_of(g, topo)
    m, t, u = m.copy(), t.copy(), u.copy()
    o.ghost_fill(m, t, u)
    want = o.rk4(m, t, u, 600.0, steps)
    _cmp_state(got, want, g.layout, TOL_ORACLE, "rk4 vs oracle")
    # frozen halo slots never change
    es = g.array("edge_src").reshape(-1, 3).T
    frozen = es < 0
    for k in range(nlayers):
        assert np.array_equal(got[2][k][frozen], u[k][frozen])


def test_full_size_trend_J8_L30():
    """The benchmark workload (J8, 30 layers): trend equals the oracle's and
    conserves mass (sum_i A_i dmu_i = 0 per layer)."""
    L = 30
    g = W.Grid(level=8, min_level=8, domain_level=2)
    topo, m, t, u = cases.ocean(g, L, seed=8)
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.trend(0, 1)
    dm, dt, du = ctx.download_state(1)
    own = g.layout.owned_cells()
    area = 1.0 / g.array("hex_inv")[own]
    for k in (0, L // 2, L - 1):
        s = np.sum(dm[k, own] * area)
        assert abs(s) < 1e-12 * np.sum(np.abs(dm[k, own] * area))
    o = oracle_of(g, topo)
    m, t, u = m.copy(), t.copy(), u.copy()
    o.ghost_fill(m, t, u)
    want = o.trend(m, t, u)
    _cmp_state((dm, dt, du), want, g.layout, TOL_ORACLE, "J8 L30 trend vs oracle")


def test_errors_are_raised():
    g = W.Grid(level=4, min_level=4, domain_level=1)
    ctx = W.Context(g, nlayers=1)
    with pytest.raises(W.WtgError, match="topography"):
        ctx.rk4(0, 1.0, 1)
    with pytest.raises(W.WtgError, match="slot"):
        ctx.trend(0, 7)


@pytest.mark.parametrize("version", ["1", "2", "3"])
def test_stage_kernel_variants_agree(version, monkeypatch):
    """Every selectable stage kernel (WTG_STAGE_KERNEL: 1 tile loop, 2 one
    thread per region cell, 3 layer pairs) gives the oracle's RK4 step to the
    last bit, including the pole and pentagon-corner tiles; 5 layers make the
    last layer pair of kernel 3 a single layer."""
    monkeypatch.setenv("WTG_STAGE_KERNEL", version)
    print(f"stage kernel {version}", flush=True)
    g = W.Grid(level=6, min_level=6, domain_level=2)
    topo, m, t, u = cases.ocean(g, 5, seed=11)
    ctx = W.Context(g, nlayers=5)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.rk4(0, 600.0, 2)
    got = ctx.download_state(0)
    ctx.trend(0, 1)
    got_trend = ctx.download_state(1)
    o = oracle_of(g, topo)
    m, t, u = m.copy(), t.copy(), u.copy()
    o.ghost_fill(m, t, u)
    want = o.rk4(m, t, u, 600.0, 2)
    _cmp_state(got, want, g.layout, 0.0, f"kernel {version} rk4 vs oracle")
    mw, tw, uw = (x.copy() for x in want)
    o.ghost_fill(mw, tw, uw)
    _cmp_state(got_trend, o.trend(mw, tw, uw), g.layout, 0.0, f"kernel {version} trend vs oracle")
