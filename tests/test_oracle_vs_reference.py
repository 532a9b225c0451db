"""CPU oracle against the reference's own trend_ml and RK4 (golden vectors).

FP64 fields must agree to 1e-12 relative (max-norm over a field, per layer);
observed agreement is ~4e-16.
"""
import numpy as np
import pytest

from conftest import golden, maxrel
from helpers import fixture_state, grid_of_fixture, oracle_of

CASES = ["tc5_J4", "ocean_J4", "ident_J4", "tc5_J5"]
TOL = 1e-12


@pytest.mark.parametrize("case", CASES)
def test_trend_matches_reference(case):
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    o = oracle_of(g, topo, f)
    o.ghost_fill(mass, temp, velo)
    dm, dt, du = o.trend(mass, temp, velo)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for k in range(int(f["nlayers"])):
        assert maxrel(dm[k, own], f["dmass"][k, own]) < TOL
        assert maxrel(dt[k, own], f["dtemp"][k, own]) < TOL if np.any(f["dtemp"][k, own]) else True
        for e in range(3):
            assert maxrel(du[k, e, own_e], f["dvelo"][k, e, own_e]) < TOL


@pytest.mark.parametrize("case", CASES)
def test_rk4_steps_match_reference(case):
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    o = oracle_of(g, topo, f)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for s in range(1, int(f["nsteps"]) + 1):
        mass, temp, velo = o.rk4(mass, temp, velo, float(f["dt"]), 1)
        for k in range(int(f["nlayers"])):
            assert maxrel(mass[k, own], f[f"mass{s}"][k, own]) < TOL
            assert maxrel(temp[k, own], f[f"temp{s}"][k, own]) < TOL if np.any(f[f"temp{s}"][k, own]) else True
            for e in range(3):
                assert maxrel(velo[k, e, own_e], f[f"velo{s}"][k, e, own_e]) < TOL


def test_c1_100_steps_match_reference():
    """The CPU-runnable case at its workload: J5 TC5, one layer, 100 RK4 steps
    at CFL 0.8 (dt 674 s), the state after steps 25, 50, 75 and 100 within
    1e-12 of the reference (observed: mass 2e-15, velocity 2e-14)"""
    f = golden("trisk_tc5_c1_J5.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    o = oracle_of(g, topo, f)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    every = int(f["every"])
    assert every * int(f["nsteps"]) == 100
    for s in range(1, int(f["nsteps"]) + 1):
        mass, temp, velo = o.rk4(mass, temp, velo, float(f["dt"]), every)
        assert maxrel(mass[0, own], f[f"mass{s}"][0, own]) < TOL
        for e in range(3):
            assert maxrel(velo[0, e, own_e], f[f"velo{s}"][0, e, own_e]) < TOL
