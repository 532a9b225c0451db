"""Laplacian diffusion and hyperdiffusion in the oracle against the reference.

Fixtures tests/golden/trisk_ocean_diff{1,2}_J4.npz come from the reference's
trend_ml and RK4 (oracle/ref_harness/dump_trisk.f90) with the physics hooks of
its test cases switched on: scalar flux (-1)^Ls visc grad(q or Laplacian q)
pedlen and velocity source (-1)^(Lr-1) (visc_divu grad divu - visc_rotu curl
rotu), on the J4 ocean case with two layers.
  diff1  Laplace_sclr = Laplace_divu = Laplace_rotu = 1, visc 2e6, 1e6, 3e6, 2e6
  diff2  all three 2 (hyperdiffusion), visc 5e17, 3e17, 6e17, 4e17

Bar: trend and two RK4 steps within 1e-13 relative (max-norm of the field
over the owned cells, per layer); observed ~1e-15.  The reference's divu on
the halo row and column i = n, j = n is each lozenge's own divergence of its
ghost fluxes (cal_divu_ml fills no ghosts after it), which differs from the
owner's value at the pentagon corner (n, n), and the halo fluxes of its
Laplacian are the owners' (update_bdry of horiz_flux); the oracle restates
both, and a planted variant with the owners' divergence everywhere moves the
velocity trend outside the bar.

Planted errors (in the oracle's parameters, on the CPU): dropping any one of
the four viscosities, or swapping the order of the scalar and velocity
operators, moves the trend far outside the bar.
"""
import numpy as np
import pytest

from conftest import golden, maxrel
from helpers import fixture_state, grid_of_fixture, oracle_of

CASES = ["ocean_diff1_J4", "ocean_diff2_J4"]
TOL = 1e-13


def load(case):
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    topo, mass, temp, velo = fixture_state(f, g)
    o = oracle_of(g, topo, f)
    o.ghost_fill(mass, temp, velo)
    lp = [int(x) for x in f["laplace"]]
    vs = [float(x) for x in f["visc"]]
    return f, g, o, (mass, temp, velo), (*lp, *vs)


def trend_errors(f, g, o, st, diff):
    dm, dt, du = o.trend(*st, diffusion=diff)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    L = int(f["nlayers"])
    em = max(maxrel(dm[k, own], f["dmass"][k, own]) for k in range(L))
    et = max(maxrel(dt[k, own], f["dtemp"][k, own]) for k in range(L))
    eu = max(maxrel(du[k, e, own_e], f["dvelo"][k, e, own_e]) for k in range(L) for e in range(3))
    return em, et, eu


@pytest.mark.parametrize("case", CASES)
def test_diffusion_trend_matches_reference(case):
    f, g, o, st, diff = load(case)
    em, et, eu = trend_errors(f, g, o, st, diff)
    assert em < TOL and et < TOL and eu < TOL, (em, et, eu)


@pytest.mark.parametrize("case", CASES)
def test_diffusion_rk4_steps_match_reference(case):
    f, g, o, st, diff = load(case)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for s in range(1, int(f["nsteps"]) + 1):
        st = o.rk4(*st, float(f["dt"]), 1, diffusion=diff)
        mass, temp, velo = st
        for k in range(int(f["nlayers"])):
            assert maxrel(mass[k, own], f[f"mass{s}"][k, own]) < TOL
            assert maxrel(temp[k, own], f[f"temp{s}"][k, own]) < TOL
            for e in range(3):
                assert maxrel(velo[k, e, own_e], f[f"velo{s}"][k, e, own_e]) < TOL


@pytest.mark.parametrize("case", CASES)
def test_planted_parameter_errors_are_caught(case):
    f, g, o, st, diff = load(case)
    Ls, Ld, Lr, vm, vt, vd, vr = diff
    other = 3 - Ls
    for bad in ((Ls, Ld, Lr, 0.0, vt, vd, vr), (Ls, Ld, Lr, vm, 0.0, vd, vr), (Ls, Ld, Lr, vm, vt, 0.0, vr),
                (Ls, Ld, Lr, vm, vt, vd, 0.0), (other, Ld, Lr, vm, vt, vd, vr), (Ls, 3 - Ld, 3 - Lr, vm, vt, vd, vr)):
        assert max(trend_errors(f, g, o, st, bad)) > 1e-9, bad


@pytest.mark.parametrize("case", CASES)
def test_owner_halo_values_are_caught(case):
    """The planted mistake of a ghost fill of divu after cal_div and of halo
    gradient fluxes computed locally (owners' values everywhere) leaves the
    bar: the reference's corner divergence is the lozenge's own"""
    f, g, o, st, diff = load(case)
    em, et, eu = trend_errors(f, g, o, st, (*diff, 1))
    assert eu > 1e-9, eu
