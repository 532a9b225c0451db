"""HIP free-surface elliptic path against the CPU oracle and the reference
(needs a GPU).

Bars: the operator, its diagonal, Jacobi sweeps, the flux divergence and the
right-hand side equal the oracle bit for bit (same evaluation order, no FMA
contraction) and the reference within 1e-12; dot products / norms within
1e-14 of the oracle (tree vs sequential summation); BiCGSTAB takes the
reference's iteration count to its 1e-9 tolerance, with the solution within
1e-10 of the reference's; at J8 the true residual of the GPU solution,
recomputed by the oracle, is within the tolerance.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden, maxrel
from helpers import grid_of_fixture
from oracle.oracle import EllipticOracle, geometry_of

pytestmark = pytest.mark.gpu


def setup(name):
    f = golden(f"elliptic_{name}.npz")
    g = grid_of_fixture(f)
    topo = np.nan_to_num(f["topo"])
    o = EllipticOracle(g.layout, geometry_of(g), topo, grav=float(f["grav"]), rho0=float(f["ref_density"]),
                       dt=float(f["dt"]), theta1=float(f["theta1"]), theta2=float(f["theta2"]))
    L = int(f["nlayers"])
    ctx = W.Context(g, nlayers=L, grav=float(f["grav"]), ref_density=float(f["ref_density"]))
    ctx.set_topography(topo)
    ctx.fs_set_params(float(f["dt"]), float(f["theta1"]), float(f["theta2"]))
    masters = g.layout.owned_cells()
    return f, g, o, ctx, masters


@pytest.mark.parametrize("name", ["J4", "J5"])
def test_operator_pieces_match_oracle_and_reference(name):
    print(f"elliptic {name}", flush=True)
    f, g, o, ctx, masters = setup(name)
    q, eo = o.node_ghost(np.nan_to_num(f["q"])), o.node_ghost(np.nan_to_num(f["eta_old"]))
    ctx.fs_upload(W.FS_ETA_OLD, eo)
    ctx.fs_upload(W.FS_ETA, q)
    ctx.elliptic_apply(W.FS_ETA, W.FS_RHS)
    Lq = ctx.fs_download(W.FS_RHS)
    want = o.apply(eo, q)
    assert np.array_equal(Lq[masters], want[masters])
    assert maxrel(Lq[masters], f["Lq"][masters]) <= 1e-12
    # halo copies written in the same launch
    src = g.array("node_src")
    halo = np.nonzero((src >= 0) & (src != np.arange(g.layout.ncells)))[0]
    assert np.array_equal(Lq[halo], Lq[src[halo]])
    ctx.elliptic_diag(W.FS_DIVF)
    assert np.array_equal(ctx.fs_download(W.FS_DIVF)[masters], o.diag(eo)[masters])
    # norms
    l2q, dpq, _ = f["red"]
    assert abs(np.sqrt(ctx.fs_dot(W.FS_ETA, W.FS_ETA)) - o.l2(q)) <= 1e-14 * l2q
    assert abs(ctx.fs_dot(W.FS_ETA, W.FS_ETA_OLD) - o.wdot(q, eo)) <= 1e-14 * abs(l2q * o.l2(eo))
    assert abs(np.sqrt(ctx.fs_dot(W.FS_ETA, W.FS_ETA)) - l2q) <= 1e-13 * l2q
    # flux divergence of the layer state, rhs, one Jacobi sweep
    mass = np.nan_to_num(f["mass"]).copy()
    velo = np.nan_to_num(f["velo"]).copy()
    temp = np.zeros_like(mass)
    ctx.upload_state(0, mass, temp, velo)
    ctx.ghost_fill(0)
    ctx.flux_divergence(0)
    dF = ctx.fs_download(W.FS_DIVF)
    o.ghost_fill(mass, temp, velo)
    dF_o = o.flux_divergence(mass, velo)
    assert np.array_equal(dF[masters], dF_o[masters])
    assert maxrel(dF[masters], f["divF"][masters]) <= 1e-12
    ctx.fs_upload(W.FS_ETA, eo)
    ctx.fs_upload(W.FS_DIVF_OLD, np.zeros(g.layout.ncells))
    ctx.rhs_elliptic()
    rhs = ctx.fs_download(W.FS_RHS)
    rhs_o = o.rhs(eo, dF_o, np.zeros(g.layout.ncells))
    assert np.array_equal(rhs[masters], rhs_o[masters])
    assert maxrel(rhs[masters], f["rhs"][masters]) <= 1e-12
    ctx.jacobi(W.FS_ETA, W.FS_RHS, float(f["wjac"]), 1)
    jac = ctx.fs_download(W.FS_ETA)
    assert np.array_equal(jac[masters], o.jacobi(eo, eo, rhs_o, float(f["wjac"]))[masters])
    assert maxrel(jac[masters], f["jac"][masters]) <= 1e-12


@pytest.mark.parametrize("name", ["J4", "J5"])
def test_bicgstab_matches_reference(name):
    f, g, o, ctx, masters = setup(name)
    eo = o.node_ghost(np.nan_to_num(f["eta_old"]))
    rhs = o.node_ghost(np.nan_to_num(f["rhs"]))
    ctx.fs_upload(W.FS_ETA_OLD, eo)
    ctx.fs_upload(W.FS_ETA, eo)
    ctx.fs_upload(W.FS_RHS, rhs)
    tol = float(f["tol"])
    iters, res = ctx.elliptic_solve(W.SOLVE_BICGSTAB, tol, 400)
    print(f"{name}: {iters} iterations, relative residual {res:.3e} (reference {int(f['iters'])})", flush=True)
    assert iters == int(f["iters"])
    assert res <= tol
    sol = ctx.fs_download(W.FS_ETA)
    assert maxrel(sol[masters], f["sol"][masters]) <= 1e-10
    # the true FP64 residual judges the solve
    r = rhs - o.apply(eo, o.node_ghost(sol))
    assert o.l2(r) / o.l2(rhs) <= tol * (1 + 1e-6)


@pytest.mark.parametrize("name", ["J4", "J5"])
def test_graph_bicgstab_equals_eager_launches(name):
    """BiCGSTAB chunks replayed from a captured graph (wtg_fs_set_graph) against
    eager launches: same iterates, residual and iteration count, bit for bit,
    over repeated solves (graph replays) and a solve stopped by max_iter."""
    print(f"graph BiCGSTAB {name}", flush=True)
    out = {}
    for graph in (True, False):
        f, g, o, ctx, masters = setup(name)
        ctx.fs_set_graph(graph)
        eo = o.node_ghost(np.nan_to_num(f["eta_old"]))
        rhs = o.node_ghost(np.nan_to_num(f["rhs"]))
        runs = []
        for tol, maxit in ((float(f["tol"]), 400), (1e-14, 3), (float(f["tol"]), 400), (1e-14, 20)):
            ctx.fs_upload(W.FS_ETA_OLD, eo)
            ctx.fs_upload(W.FS_ETA, eo)
            ctx.fs_upload(W.FS_RHS, rhs)
            iters, res = ctx.elliptic_solve(W.SOLVE_BICGSTAB, tol, maxit)
            runs.append((iters, res, ctx.fs_download(W.FS_ETA)))
        out[graph] = runs
    for (i1, r1, s1), (i2, r2, s2) in zip(out[True], out[False]):
        assert i1 == i2 and r1 == r2
        assert np.array_equal(s1, s2)
    assert out[True][1][0] == 4  # max_iter + 1: not converged


@pytest.mark.parametrize("name", ["J5"])
def test_solves_in_sequence_equal_fresh_solves(name):
    """The first chunk of gated iterations follows the previous solve on the
    context (at the same tolerance): a sequence of solves at different
    tolerances and iteration limits on one context gives every solve's
    iterates, residual and count bit for bit as a fresh context does."""
    print(f"BiCGSTAB solves in sequence {name}", flush=True)
    seq = ((1e-2, 400), (1e-12, 400), (1e-2, 400), (1e-2, 2), (1e-12, 400), (1e-6, 400))
    f, g, o, ctx, masters = setup(name)
    eo = o.node_ghost(np.nan_to_num(f["eta_old"]))
    rhs = o.node_ghost(np.nan_to_num(f["rhs"]))

    def solve(c, tol, maxit):
        c.fs_upload(W.FS_ETA_OLD, eo)
        c.fs_upload(W.FS_ETA, eo)
        c.fs_upload(W.FS_RHS, rhs)
        iters, res = c.elliptic_solve(W.SOLVE_BICGSTAB, tol, maxit)
        return iters, res, c.fs_download(W.FS_ETA)

    for tol, maxit in seq:
        i1, r1, s1 = solve(ctx, tol, maxit)
        fresh = setup(name)[3]
        i2, r2, s2 = solve(fresh, tol, maxit)
        fresh.close()
        print(f"  tol {tol:.0e} maxit {maxit}: {i1} iterations, residual {r1:.3e}", flush=True)
        assert i1 == i2 and r1 == r2 and np.array_equal(s1, s2)
        assert (r1 <= tol) == (i1 <= maxit)


def test_jacobi_sweeps_equal_oracle_J4():
    f, g, o, ctx, masters = setup("J4")
    eo = o.node_ghost(np.nan_to_num(f["eta_old"]))
    rhs = o.node_ghost(np.nan_to_num(f["rhs"]))
    ctx.fs_upload(W.FS_ETA_OLD, eo)
    ctx.fs_upload(W.FS_ETA, eo)
    ctx.fs_upload(W.FS_RHS, rhs)
    ctx.jacobi(W.FS_ETA, W.FS_RHS, 0.9, 7)
    u = eo
    for _ in range(7):
        u = o.jacobi(eo, u, rhs, 0.9)
    assert np.array_equal(ctx.fs_download(W.FS_ETA)[masters], u[masters])
    # the Jacobi solver drives the residual to its tolerance
    ctx.fs_upload(W.FS_ETA, eo)
    iters, res = ctx.elliptic_solve(W.SOLVE_JACOBI, 1e-8, 500)
    assert res <= 1e-8 and iters <= 500
    sol = ctx.fs_download(W.FS_ETA)
    r = rhs - o.apply(eo, o.node_ghost(sol))
    assert o.l2(r) / o.l2(rhs) <= 1e-8 * (1 + 1e-6)


def test_full_size_J8_solve():
    """The J8 ocean free surface (655 362 nodes): operator bit-exact against
    the oracle on random data, and a BiCGSTAB solve to 1e-9 whose residual
    the oracle confirms."""
    print("J8 elliptic", flush=True)
    g = W.Grid(level=8, min_level=8, domain_level=2)
    from wavetrisk_hydrostatic_amd import cases
    topo, mass, temp, velo = cases.ocean(g, 2, seed=3)
    nc = g.layout.ncells
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=300.0)
    ctx = W.Context(g, nlayers=2)
    ctx.set_topography(topo)
    ctx.fs_set_params(300.0, 1.0, 1.0)
    rng = np.random.default_rng(8)
    masters = g.layout.owned_cells()
    q = o.node_ghost(rng.standard_normal(nc))
    eo = o.node_ghost(0.1 * rng.standard_normal(nc))
    ctx.fs_upload(W.FS_ETA_OLD, eo)
    ctx.fs_upload(W.FS_ETA, q)
    ctx.elliptic_apply(W.FS_ETA, W.FS_DIVF)
    assert np.array_equal(ctx.fs_download(W.FS_DIVF)[masters], o.apply(eo, q)[masters])
    rhs = o.node_ghost(-eo + 0.01 * rng.standard_normal(nc))
    ctx.fs_upload(W.FS_RHS, rhs)
    ctx.fs_upload(W.FS_ETA, eo)
    iters, res = ctx.elliptic_solve(W.SOLVE_BICGSTAB, 1e-9, 400)
    print(f"J8: {iters} iterations, residual {res:.3e}", flush=True)
    assert res <= 1e-9 and iters <= 400
    sol = ctx.fs_download(W.FS_ETA)
    r = rhs - o.apply(eo, o.node_ghost(sol))
    assert o.l2(r) / o.l2(rhs) <= 1e-9 * (1 + 1e-6)


def test_elliptic_errors():
    g = W.Grid(level=4, min_level=4, domain_level=1)
    ctx = W.Context(g, nlayers=1)
    with pytest.raises(W.WtgError, match="topography"):
        ctx.fs_upload(W.FS_ETA, np.zeros(g.layout.ncells))
    ctx.set_topography(np.full(g.layout.ncells, -1000.0))
    with pytest.raises(W.WtgError, match="wtg_fs_set_params"):
        ctx.elliptic_apply(W.FS_ETA, W.FS_RHS)
    ctx.fs_set_params(100.0)
    with pytest.raises(W.WtgError, match="field"):
        ctx.elliptic_apply(W.FS_ETA, 9)
    with pytest.raises(W.WtgError, match="differ"):
        ctx.elliptic_apply(W.FS_ETA, W.FS_ETA)
    with pytest.raises(W.WtgError, match="method"):
        ctx.elliptic_solve(7, 1e-9, 10)
