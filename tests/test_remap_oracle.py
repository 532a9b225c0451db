"""Vertical remap of the Lagrangian layers (remap_vertical_coordinates) in the
oracle against the reference.

Fixtures tests/golden/trisk_ocean_remap{1,0}_J4.npz come from the
reference's remap_vertical_coordinates (remap.f90:18-75, incompressible
branch 147-257) run by the harness (oracle/ref_harness/dump_trisk.f90) after
one RK4 step of the J4 ocean with 6 layers: remap_type "1" (piecewise linear)
and "2PPM" (Colella-Woodward piecewise parabolic) onto Chebyshev hybrid
coordinates and remap_type "0" (piecewise constant) onto uniform ones (b_vert
as the test cases' initialize_a_b_vert, a_vert = 1 - b_vert).  The
reference's default, "PPR" (Engwirda's piecewise-polynomial library), is not
restated.

Bars: the oracle equals the reference to 1e-15 relative (observed: bit for
bit); closed-form properties of the finite-volume remap hold to rounding:
column mass and column buoyancy content (sum of the mass-weighted buoyancy)
are conserved, and remapping a remapped state onto the same coordinates
changes nothing.  Planted errors (in the oracle's inputs, on the CPU): the
other scheme, or the target coordinates moved by 1e-6 of the depth, leave
the bar.
"""
import numpy as np
import pytest

from conftest import golden, maxrel
from helpers import grid_of_fixture, oracle_of

CASES = ["ocean_remap1_J4", "ocean_remap0_J4", "ocean_remap2ppm_J4"]


def load(case):
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    o = oracle_of(g, np.nan_to_num(f["topo"]), f)
    ns = int(f["nsteps"])
    m, t, u = (np.nan_to_num(f[f"{v}{ns}"]).copy() for v in ("mass", "temp", "velo"))
    o.ghost_fill(m, t, u)
    return f, g, o, (m, t, u), int(str(f["remap_scheme"]))


def errors(f, g, got):
    oc, oe = g.layout.owned_cells(), g.layout.owned_cells(with_poles=False)
    m, t, u = got
    L = m.shape[0]
    return (max(maxrel(m[k, oc], f["remapped_mass"][k, oc]) for k in range(L)),
            max(maxrel(t[k, oc], f["remapped_temp"][k, oc]) for k in range(L)),
            max(maxrel(u[k, :, oe], f["remapped_velo"][k, :, oe]) for k in range(L)))


@pytest.mark.parametrize("case", CASES)
def test_remap_matches_reference(case):
    f, g, o, st, scheme = load(case)
    got = o.remap(*st, scheme, f["a_vert"], f["b_vert"])
    em, et, eu = errors(f, g, got)
    assert em <= 1e-15 and et <= 1e-15 and eu <= 1e-15, (em, et, eu)
    # the remap is not a no-op on this state
    oc = g.layout.owned_cells()
    assert maxrel(got[0][:, oc], st[0][:, oc]) > 1e-3


@pytest.mark.parametrize("case", CASES)
def test_remap_conserves_columns_and_is_idempotent(case):
    f, g, o, st, scheme = load(case)
    m, t, u = o.remap(*st, scheme, f["a_vert"], f["b_vert"])
    oc = g.layout.owned_cells()
    m0, t0 = st[0][:, oc], st[1][:, oc]
    assert np.max(np.abs(m[:, oc].sum(0) - m0.sum(0)) / m0.sum(0)) <= 1e-14
    assert np.max(np.abs(t[:, oc].sum(0) - t0.sum(0))) / np.max(np.abs(t0.sum(0))) <= 1e-12
    m2, t2, u2 = o.remap(m, t, u, scheme, f["a_vert"], f["b_vert"])
    oe = g.layout.owned_cells(with_poles=False)
    assert maxrel(m2[:, oc], m[:, oc]) <= 1e-14 and maxrel(t2[:, oc], t[:, oc]) <= 1e-12
    assert maxrel(u2[:, :, oe], u[:, :, oe]) <= 1e-12


@pytest.mark.parametrize("case", CASES)
def test_planted_errors_are_caught(case):
    f, g, o, st, scheme = load(case)
    other = o.remap(*st, (scheme + 1) % 3, f["a_vert"], f["b_vert"])
    assert max(errors(f, g, other)) > 1e-9
    b = f["b_vert"].copy()
    b[1:-1] += 1e-6  # the inner interfaces moved by 1e-6 of the column depth
    moved = o.remap(*st, scheme, 1.0 - b, b)
    assert max(errors(f, g, moved)) > 1e-9
