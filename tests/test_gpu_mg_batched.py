"""Member-batched multigrid on a hierarchy of ensemble contexts (needs a GPU).

Several right-hand sides (each with its own bathymetry and old free surface)
solved in the same launches -- plain V-cycles, learned cycles, and the
split-explicit step of an ensemble with its free surface on the hierarchy --
against each member solved alone on its own single-member hierarchy.

Bar: bit-identical solutions and the same cycle count per member (a
converged member's fields are left alone, the members' reductions are
separate, and the network's rows do not interact), at J4..J6 with three
members; the members converge after different numbers of cycles, so the
gating is exercised.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import learned as LC

pytestmark = pytest.mark.gpu

SEEDS = [3, 17, 29]


@pytest.fixture(scope="module")
def grids():
    return LC.hierarchy_grids(fine=6, levels=3)


def _alone(grids, method, p, mlp_seed=None):
    out = []
    for s in SEEDS:
        prob = LC.Problem(grids, s)
        mlp = None
        try:
            if mlp_seed is not None:
                mlp = W.MLP(max_batch=prob.h.npatches * len(SEEDS), seed=mlp_seed)
                prob.h.set_learned(mlp)
            it, res, _ = prob.solve(method, p)
            out.append((int(it[-1]), prob.top.fs_download(W.FS_ETA)))
        finally:
            if mlp is not None:
                prob.h.set_learned(None)
                mlp.close()
            prob.close()
    return out


@pytest.mark.parametrize("method", ["vcycle", "learned"])
def test_batched_cycles_equal_single_solves(grids, method):
    print(f"member-batched {method} cycles, J4..J6, {len(SEEDS)} members", flush=True)
    meth = W.SOLVE_VCYCLE if method == "vcycle" else W.SOLVE_LEARNED
    p = LC.vcycle_params(1e-10, 40)
    mlp_seed = 7 if method == "learned" else None
    alone = _alone(grids, meth, p, mlp_seed)
    prob = LC.Problem(grids, SEEDS)
    mlp = None
    try:
        if mlp_seed is not None:
            mlp = W.MLP(max_batch=prob.h.npatches * len(SEEDS), seed=mlp_seed)
            prob.h.set_learned(mlp)
        it, res, _ = prob.solve(meth, p)
        cyc, mres = prob.h.member_results()
        u = prob.top.fs_download(W.FS_ETA)
    finally:
        if mlp is not None:
            prob.h.set_learned(None)
            mlp.close()
        prob.close()
    print(f"  cycles per member {list(cyc)} (alone {[a[0] for a in alone]}), residuals {mres}", flush=True)
    assert list(cyc) == [a[0] for a in alone]
    assert it[-1] == max(cyc) and np.all(mres <= 1e-10)
    for k, (_, ua) in enumerate(alone):
        assert np.array_equal(u[k], ua), f"member {k} differs"


def test_ensemble_split_step_on_batched_hierarchy():
    """wtg_mg_rk4_split on an ensemble: the members' trends in one launch and
    their free surfaces solved together by V-cycles on the hierarchy -- each
    member equal to its own single-member run, bit for bit"""
    print("ensemble split step with member-batched V-cycles, J4..J6", flush=True)
    from wavetrisk_hydrostatic_amd import cases
    grids = [W.Grid(level=l, min_level=4, domain_level=1) for l in (4, 5, 6)]
    L, dt, M = 3, 1800.0, 2
    draws = [[cases.ocean(g, L, seed=40 + k) for g in grids] for k in range(M)]
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-11, coarse_iter=400, max_vcycle=60)

    def run(members):
        ctxs = []
        for lv, g in enumerate(grids):
            nl = L if lv == 2 else 1
            c = W.Context(g, nlayers=nl, members=len(members))
            topo = np.stack([draws[k][lv][0] for k in members]) if len(members) > 1 else draws[members[0]][lv][0]
            c.set_topography(topo)
            c.fs_set_params(dt, 0.6, 0.8)
            ctxs.append(c)
        top = ctxs[-1]
        m = np.concatenate([draws[k][2][1] for k in members])
        t = np.concatenate([draws[k][2][2] for k in members])
        u = np.concatenate([draws[k][2][3] for k in members])
        eta = [draws[k][2][1].sum(0) / 1028.0 + draws[k][2][0] for k in members]
        top.upload_state(0, m, t, u)
        top.fs_upload(W.FS_ETA, np.stack(eta) if len(members) > 1 else eta[0])
        h = W.Hierarchy(ctxs)
        try:
            it = h.rk4_split(0, 2, W.SOLVE_VCYCLE, p)
        finally:
            h.close()
        res = top.download_state(0), top.fs_download(W.FS_ETA), it
        for c in ctxs:
            c.close()
        return res

    (mb, tb, ub), eb, itb = run(list(range(M)))
    own, owne = grids[-1].layout.owned_cells(), grids[-1].layout.owned_edge_cells()
    its = []
    for k in range(M):
        (ma, ta, ua), ea, ita = run([k])
        its.append(ita)
        sl = slice(k * L, (k + 1) * L)
        assert np.array_equal(mb[sl][:, own], ma[:, own]) and np.array_equal(tb[sl][:, own], ta[:, own])
        assert np.array_equal(ub[sl][:, :, owne], ua[:, :, owne])
        assert np.array_equal(eb[k][own], ea[own])
    assert itb == max(its)


def test_ensemble_j8_l30_learned_split_step():
    """Config 5's shape: two members of the J8 L30 ocean, each with its own
    bathymetry, one split-explicit step with the free surface solved by
    learned cycles on the member-batched J6..J8 hierarchy (any network).
    Each member equals its own single-member run bit for bit; member 0
    against the oracle's split step (single-level BiCGSTAB to 1e-12) within
    the solver tolerance's reach (1e-8)."""
    print("J8 L30 ensemble of 2, learned split step on J6..J8", flush=True)
    from wavetrisk_hydrostatic_amd import cases
    from conftest import maxrel
    from oracle.oracle import EllipticOracle, geometry_of
    grids = [W.Grid(level=l, min_level=6, domain_level=2) for l in (6, 7, 8)]
    L, dt, M = 30, 900.0, 2
    draws = [[cases.ocean(g, L if lv == 2 else 1, seed=60 + k) for lv, g in enumerate(grids)] for k in range(M)]
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-10, coarse_iter=400, max_vcycle=60)

    def run(members):
        ctxs = []
        for lv, g in enumerate(grids):
            c = W.Context(g, nlayers=L if lv == 2 else 1, members=len(members))
            topo = np.stack([draws[k][lv][0] for k in members]) if len(members) > 1 else draws[members[0]][lv][0]
            c.set_topography(topo)
            c.fs_set_params(dt, 0.6, 0.8)
            ctxs.append(c)
        top = ctxs[-1]
        top.upload_state(0, *(np.concatenate([draws[k][2][v] for k in members]) for v in (1, 2, 3)))
        eta = [draws[k][2][1].sum(0) / 1028.0 + draws[k][2][0] for k in members]
        top.fs_upload(W.FS_ETA, np.stack(eta) if len(members) > 1 else eta[0])
        h = W.Hierarchy(ctxs)
        mlp = W.MLP(max_batch=h.npatches * len(members), seed=9)
        try:
            h.set_learned(mlp)
            it = h.rk4_split(0, 1, W.SOLVE_LEARNED, p)
        finally:
            h.set_learned(None)
            h.close()
            mlp.close()
        res = top.download_state(0), top.fs_download(W.FS_ETA), it
        for c in ctxs:
            c.close()
        return res

    (mb, tb, ub), eb, itb = run([0, 1])
    g = grids[-1]
    own, owne = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for k in range(M):
        (ma, ta, ua), ea, _ = run([k])
        sl = slice(k * L, (k + 1) * L)
        assert np.array_equal(mb[sl][:, own], ma[:, own]) and np.array_equal(ub[sl][:, :, owne], ua[:, :, owne])
        assert np.array_equal(eb[k][own], ea[own])
    topo, m0, t0, u0 = draws[0][2]
    o = EllipticOracle(g.layout, geometry_of(g), topo, dt=dt, theta1=0.6, theta2=0.8)
    m0, t0, u0 = m0.copy(), t0.copy(), u0.copy()
    o.ghost_fill(m0, t0, u0)
    mo, _, uo, eo, _ = o.rk4_split(m0, t0, u0, o.node_ghost(m0.sum(0) / 1028.0 + topo), tol=1e-12, maxit=2000)
    print(f"  member 0 vs oracle: eta {maxrel(eb[0][own], eo[own]):.2e}, "
          f"velocity {maxrel(ub[:L][:, :, owne], uo[:, :, owne]):.2e}", flush=True)
    assert maxrel(eb[0][own], eo[own]) <= 1e-8
    assert maxrel(mb[:L][:, own], mo[:, own]) <= 1e-8
    assert maxrel(ub[:L][:, :, owne], uo[:, :, owne]) <= 1e-8
