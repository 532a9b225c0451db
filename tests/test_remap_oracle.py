"""Vertical remap of the Lagrangian layers (remap_vertical_coordinates) in the
oracle against the reference.

Fixtures tests/golden/trisk_ocean_remap*_J4.npz come from the reference's
remap_vertical_coordinates (remap.f90:18-75, incompressible branch 147-257)
run by the harness (oracle/ref_harness/dump_trisk.f90) after one RK4 step of
the J4 ocean with 6 layers: remap_type "1" (piecewise linear), "2PPM"
(Colella-Woodward piecewise parabolic), "PPR" (Engwirda's library at
order 2, the reference's default), "2S" (parabolic spline), "2W" (parabolic
WENO) and "4" (parabolic WENO with power-law reconciliation) onto Chebyshev
hybrid coordinates, "0" (piecewise constant) onto uniform ones (b_vert as the
test cases' initialize_a_b_vert, a_vert = 1 - b_vert).  remap4's three Pade
constants are single-precision literals in the reference: written as double
literals, 15,371 buoyancy and 30,520 velocity values of the "4" fixture move
(up to 5e-14), so the fixture pins that reading.  Two more PPR fixtures take the
branches the first leaves: "ppr_uniform" remaps old layers that a
piecewise-constant remap first made evenly spaced (rmap1d's uniform-spacing
branch: plsc slopes, the closed-form cubic edge values), "ppr_l3" has three
layers (p1e's two-point edge values).

Bars: the oracle equals the reference to 1e-15 relative (observed: bit for
bit); closed-form properties of the finite-volume remap hold to rounding:
column mass and column buoyancy content (sum of the mass-weighted buoyancy)
are conserved, and remapping a remapped state onto the same coordinates
changes nothing; PPR remaps the cell means of a line onto any cells as the
exact means of that line on the inner cells and of the end cells' own means
on the end cells (its cubic edge fits and loose end fits reproduce a line,
and the mono limiter flattens the two end cells).
Planted errors (in the oracle's inputs, on the CPU): another scheme, or the
target coordinates moved by 1e-6 of the depth, leave the bar.
"""
import numpy as np
import pytest

from conftest import golden, maxrel
from helpers import grid_of_fixture, oracle_of
from oracle.oracle import remap_column

CASES = ["ocean_remap1_J4", "ocean_remap0_J4", "ocean_remap2ppm_J4", "ocean_remapppr_J4", "ocean_remapppr_uniform_J4",
         "ocean_remapppr_l3_J4", "ocean_remap2s_J4", "ocean_remap2w_J4", "ocean_remap4_J4"]


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
    other = o.remap(*st, (scheme + 1) % 7, f["a_vert"], f["b_vert"])
    assert max(errors(f, g, other)) > 1e-9
    b = f["b_vert"].copy()
    b[1:-1] += 1e-6  # the inner interfaces moved by 1e-6 of the column depth
    moved = o.remap(*st, scheme, 1.0 - b, b)
    assert max(errors(f, g, moved)) > 1e-9


def _line_with_flat_ends(a, b, zo, zn):
    """Exact cell means on zn of the profile PPR reconstructs from the cell
    means of a + b z on zo (four or more cells): the line on the inner cells,
    and on each end cell its own mean (the mono limiter flattens the end cells,
    whose outer neighbour is themselves)"""
    N = zo.size - 1
    lo_f = a + b * (zo[0] + zo[1]) / 2
    hi_f = a + b * (zo[N - 1] + zo[N]) / 2
    out = np.empty(N)
    for k in range(N):
        x0, x1 = zn[k], zn[k + 1]
        tot = 0.0
        for (p, q, f) in ((zo[0], zo[1], lo_f), (zo[N - 1], zo[N], hi_f)):
            tot += f * max(0.0, min(q, x1) - max(p, x0))
        p, q = max(zo[1], x0), min(zo[N - 1], x1)
        if q > p:
            tot += a * (q - p) + b * (q * q - p * p) / 2
        out[k] = tot / (x1 - x0)
    return out


# two or three layers take p1e, whose end values are the end cells' values
@pytest.mark.parametrize("N", [4, 5, 9, 30])
@pytest.mark.parametrize("spacing", ["uniform", "random"])
def test_ppr_reproduces_linear_profiles(N, spacing):
    rng = np.random.default_rng(N)
    zo = np.linspace(-4000.0, 3.0, N + 1)
    if spacing == "random":
        zo[1:-1] += rng.uniform(-0.3, 0.3, N - 1) * (zo[1] - zo[0])
    zn = np.sort(np.concatenate([[zo[0]], rng.uniform(zo[0], zo[-1], N - 1), [zo[-1]]]))
    a, b = 2.5, 3e-4
    line = lambda z: a + b * (z[1:] + z[:-1]) / 2  # cell means of a + b z
    got = remap_column(3, line(zo), zo, zn)
    assert np.max(np.abs(got - _line_with_flat_ends(a, b, zo, zn))) <= 1e-12 * a
    # the same remap of a kinked profile is not the line's
    kink = line(zo) + np.where(np.arange(N) == N // 2, 0.5, 0.0)
    assert np.max(np.abs(remap_column(3, kink, zo, zn) - _line_with_flat_ends(a, b, zo, zn))) > 1e-3


def test_ppr_column_conserves_and_is_monotone():
    rng = np.random.default_rng(7)
    for N in (4, 6, 20, 40):
        zo = np.cumsum(np.concatenate([[-3000.0], rng.uniform(20, 200, N)]))
        zn = np.linspace(zo[0], zo[-1], N + 1)
        v = np.cumsum(rng.uniform(0, 1, N))  # monotone data
        got = remap_column(3, v, zo, zn)
        assert abs(np.dot(got, np.diff(zn)) - np.dot(v, np.diff(zo))) <= 1e-13 * np.dot(np.abs(v), np.diff(zo))
        assert np.all(np.diff(got) >= -1e-12 * np.max(np.abs(v)))  # the mono limiter keeps it monotone
        assert got.min() >= v.min() - 1e-12 and got.max() <= v.max() + 1e-12
