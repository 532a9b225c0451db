"""The CPU oracle of the free-surface elliptic path against the reference.

Fixtures tests/golden/elliptic_J{4,5}.npz come from the reference's own
routines (oracle/ref_harness/dump_elliptic.f90): elliptic_lo, its diagonal,
a Jacobi sweep, flux_divergence, rhs_elliptic, l2/dp and a BiCGSTAB solve.

Bars: fields within 1e-12 relative of the reference (max-norm over owned
nodes and poles; the operator itself is bit-exact away from the cells next
to the poles, where the reference's per-lozenge pole copies and the grid's
last-bit geometry differ); l2 / dp within 1e-13 (summation order); BiCGSTAB
reaches the reference's tolerance in the same number of iterations with a
solution within 1e-10.

The mathematics pins what the reference leaves open: the operator minus
-I is symmetric and negative semidefinite in the area-weighted inner product,
annihilates constants, and the flux divergence conserves mass.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden, maxrel
from helpers import grid_of_fixture
from oracle.oracle import EllipticOracle, geometry_of

CASES = ["J4", "J5"]
TOL_REF = 1e-12


def setup(name):
    f = golden(f"elliptic_{name}.npz")
    g = grid_of_fixture(f)
    o = EllipticOracle(g.layout, geometry_of(g), np.nan_to_num(f["topo"]), grav=float(f["grav"]),
                       rho0=float(f["ref_density"]), dt=float(f["dt"]), theta1=float(f["theta1"]),
                       theta2=float(f["theta2"]))
    masters = g.layout.owned_cells()  # owned nodes and the two pole masters
    field = lambda k: o.node_ghost(np.nan_to_num(f[k]))
    return f, g, o, masters, field


@pytest.mark.parametrize("name", CASES)
def test_reduction_weights_are_the_reference_sweep(name):
    f, g, o, masters, field = setup(name)
    node_src = g.array("node_src")
    w = np.zeros(g.layout.ncells)
    for cell, count in f["visits"]:
        owner = cell if cell in (g.layout.south_pole, g.layout.north_pole) or node_src[cell] == cell else node_src[cell]
        assert owner >= 0
        w[owner] += count
    assert np.array_equal(w, g.array("red_weight"))


@pytest.mark.parametrize("name", CASES)
def test_norms_match_reference(name):
    f, g, o, masters, field = setup(name)
    q, eo = field("q"), field("eta_old")
    l2q, dpq, l2e = f["red"]
    assert abs(o.l2(q) - l2q) <= 1e-13 * l2q
    assert abs(o.wdot(q, eo) - dpq) <= 1e-13 * abs(l2q * l2e)
    assert abs(o.l2(eo) - l2e) <= 1e-13 * l2e


@pytest.mark.parametrize("name", CASES)
def test_operator_and_diagonal_match_reference(name):
    f, g, o, masters, field = setup(name)
    q, eo = field("q"), field("eta_old")
    Lq = o.apply(eo, q)
    assert maxrel(Lq[masters], f["Lq"][masters]) <= TOL_REF
    assert np.mean(Lq[masters] == f["Lq"][masters]) > 0.99  # bit-exact away from the pole neighbourhoods
    assert maxrel(o.diag(eo)[masters], f["diag"][masters]) <= TOL_REF


@pytest.mark.parametrize("name", CASES)
def test_rhs_jacobi_and_bicgstab_match_reference(name):
    f, g, o, masters, field = setup(name)
    L = int(f["nlayers"])
    nc = g.layout.ncells
    mass = np.nan_to_num(f["mass"]).copy()
    velo = np.nan_to_num(f["velo"]).copy()
    temp = np.zeros_like(mass)
    o.ghost_fill(mass, temp, velo)
    divF = o.flux_divergence(mass, velo)
    assert maxrel(divF[masters], f["divF"][masters]) <= TOL_REF
    eo = field("eta_old")
    rhs = o.rhs(eo, divF, np.zeros(nc))
    assert maxrel(rhs[masters], f["rhs"][masters]) <= TOL_REF
    jac = o.jacobi(eo, eo, rhs, float(f["wjac"]))
    assert maxrel(jac[masters], f["jac"][masters]) <= TOL_REF
    sol, iters, err = o.bicgstab(eo, eo, rhs, float(f["tol"]), 400)
    assert iters == int(f["iters"])
    assert err <= float(f["tol"])
    assert abs(err - f["err"][0]) <= 1e-3 * f["err"][0]
    assert maxrel(sol[masters], f["sol"][masters]) <= 1e-10
    _ = L


def test_operator_properties_J5():
    """sum_i A_i p_i (L q)_i is symmetric in p, q; q.(L+I)q <= 0; L c = -c."""
    f, g, o, masters, field = setup("J5")
    rng = np.random.default_rng(5)
    nc = g.layout.ncells
    area = np.zeros(nc)
    area[masters] = 1.0 / g.array("hex_inv")[masters]
    eo = field("eta_old")
    p = o.node_ghost(np.where(np.isin(np.arange(nc), masters), rng.standard_normal(nc), 0.0))
    q = o.node_ghost(np.where(np.isin(np.arange(nc), masters), rng.standard_normal(nc), 0.0))
    Lp, Lq = o.apply(eo, p), o.apply(eo, q)
    a = np.sum(area[masters] * p[masters] * Lq[masters])
    b = np.sum(area[masters] * q[masters] * Lp[masters])
    assert abs(a - b) <= 1e-11 * np.sum(area[masters] * np.abs(p[masters] * Lq[masters]))
    D = Lq + q  # theta1 theta2 dt^2 div(c^2 grad q)
    assert np.sum(area[masters] * q[masters] * D[masters]) < 0
    c = o.node_ghost(np.full(nc, 0.37))
    Lc = o.apply(eo, c)
    assert np.max(np.abs(Lc[masters] + 0.37)) <= 1e-15


def test_planted_errors_are_caught():
    """The symmetry check is not vacuous: restated in numpy, the operator
    passes it, and fails it once a plausible mistake is planted -- the SW
    neighbour taken one cell off, or the S edge's flux dropped from the
    divergence (fields supported away from the lozenge edges, so the
    interior restatement sees every flux)."""
    f, g, o, masters, field = setup("J4")
    rng = np.random.default_rng(1)
    nc, S, n = g.layout.ncells, g.layout.S, g.layout.n
    lz, i, j = g.layout.coords()
    inner = np.nonzero((i >= 2) & (i < n - 2) & (j >= 2) & (j < n - 2))[0]
    support = (i >= 3) & (i < n - 3) & (j >= 3) & (j < n - 3)
    p = np.where(support, rng.standard_normal(nc), 0.0)
    q = np.where(support, rng.standard_normal(nc), 0.0)
    hinv = g.array("hex_inv")
    area = np.zeros(nc)
    area[inner] = 1.0 / hinv[inner]
    lenf, pedf = g.array("len"), g.array("pedlen")
    eo = field("eta_old")
    csq = float(f["grav"]) * (np.abs(o._keep["topo"]) + eo)

    def op(qq, plant):
        def flux(x, e):
            nb = x + (1 if e == 0 else (S + 1 if e == 1 else S))
            d = qq[x] - qq[nb] if e == 1 else qq[nb] - qq[x]
            return 0.5 * (csq[x] + csq[nb]) * d / lenf[e, x] * pedf[e, x]
        out = np.zeros(nc)
        for c in inner:
            sw = c - S if plant == "sw" else c - S - 1
            fs = 0.0 if plant == "drop" else flux(c - S, 2)
            out[c] = (flux(c, 0) - flux(c - 1, 0) + flux(sw, 1) - flux(c, 1) + flux(c, 2) - fs) * hinv[c]
        return out

    def asym(plant):
        Dp, Dq = op(p, plant), op(q, plant)
        a, b = np.sum(area[inner] * p[inner] * Dq[inner]), np.sum(area[inner] * q[inner] * Dp[inner])
        return abs(a - b) / np.sum(area[inner] * np.abs(p[inner] * Dq[inner]))

    assert asym(None) <= 1e-12
    assert asym("sw") > 1e-3
    assert asym("drop") > 1e-3
