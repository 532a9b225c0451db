"""Batched ensemble (config 5's shape, one launch per operator for all
members) against single-member contexts and the CPU oracle (needs a GPU).

Bar: every member of an ensemble RK4 step equals the same member stepped
alone bit for bit, and the oracle (the stage kernel is bit-exact against it).
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from oracle.oracle import Oracle, geometry_of

pytestmark = pytest.mark.gpu


def test_ensemble_members_equal_single_runs():
    print("ensemble J5 L4 x 3", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, M, dt = 4, 3, 300.0
    draws = [cases.ocean(g, L, seed=10 + m) for m in range(M)]
    ens = W.Context(g, nlayers=L, members=M)
    ens.set_topography(np.stack([d[0] for d in draws]))
    ens.upload_state(0, *(np.concatenate([d[k] for d in draws]) for k in (1, 2, 3)))
    ens.rk4(0, dt, 2)
    em, et, eu = ens.download_state(0)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    for m, (topo, mass, temp, velo) in enumerate(draws):
        one = W.Context(g, nlayers=L)
        one.set_topography(topo)
        one.upload_state(0, mass, temp, velo)
        one.rk4(0, dt, 2)
        sm, st, su = one.download_state(0)
        sl = slice(m * L, (m + 1) * L)
        assert np.array_equal(em[sl][:, own], sm[:, own])
        assert np.array_equal(et[sl][:, own], st[:, own])
        assert np.array_equal(eu[sl][:, :, own_e], su[:, :, own_e])
        if m == M - 1:
            o = Oracle(g.layout, geometry_of(g), topo)
            mm, tt, uu = mass.copy(), temp.copy(), velo.copy()
            o.ghost_fill(mm, tt, uu)
            wm, _, wu = o.rk4(mm, tt, uu, dt, 2)
            assert np.array_equal(sm[:, own], wm[:, own])
            assert np.array_equal(su[:, :, own_e], wu[:, :, own_e])


def test_ensemble_free_surface_restrictions():
    """An ensemble's free surface runs through the split step or a
    member-batched hierarchy: the single-field solver entry points refuse
    it; a hierarchy needs the same member count on every level and solves by
    V-cycles or learned cycles only (FMG and SRJ follow one right-hand side's
    per-level decisions)."""
    g = W.Grid(level=4, min_level=4, domain_level=1)
    g5 = W.Grid(level=5, min_level=4, domain_level=1)
    ens = W.Context(g, nlayers=2, members=2)
    ens.set_topography(np.full(g.layout.ncells, -1000.0))
    ens.fs_set_params(100.0)
    with pytest.raises(W.WtgError, match="ensemble"):
        ens.elliptic_apply(W.FS_ETA, W.FS_RHS)
    three = W.Context(g5, nlayers=1, members=3)
    three.set_topography(np.full(g5.layout.ncells, -1000.0))
    three.fs_set_params(100.0)
    with pytest.raises(W.WtgError, match="members"):
        W.Hierarchy([ens, three])
    h = W.Hierarchy([ens])
    try:
        with pytest.raises(W.WtgError, match="ensemble hierarchy"):
            h.solve(W.SOLVE_FMG)
        with pytest.raises(W.WtgError, match="ensemble hierarchy"):
            h.set_precision(W.PRECISION_FP32_BULK)
    finally:
        h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_ensemble_split_steps_equal_single_runs(precision):
    """Split-explicit steps (RK4_split with the implicit free surface) of an
    ensemble: one trend launch for all members, each member's free surface
    solved on its own view.  Every member equals the same member stepped
    alone bit for bit (state and free surface); the reported iteration count
    is the largest of the members'."""
    print(f"ensemble split steps J5 L3 x 3 ({precision})", flush=True)
    prec = W.PRECISION_FP64 if precision == "fp64" else W.PRECISION_FP32_BULK
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, M, dt = 3, 3, 1800.0
    draws = [cases.ocean(g, L, seed=30 + m) for m in range(M)]
    etas = [d[1].sum(0) / 1028.0 + d[0] for d in draws]
    ens = W.Context(g, nlayers=L, members=M)
    ens.set_topography(np.stack([d[0] for d in draws]))
    ens.upload_state(0, *(np.concatenate([d[k] for d in draws]) for k in (1, 2, 3)))
    ens.fs_set_params(dt, 0.6, 0.8)
    ens.fs_upload(W.FS_ETA, np.stack(etas))
    ens.set_precision(prec)
    ens.set_stage_shape(W.SHAPE_LAYER_PAIRS)
    it_e = ens.rk4_split(0, 2, 1e-11, 400)
    em, et, eu = ens.download_state(0)
    ee = ens.fs_download(W.FS_ETA)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    its = []
    for m, (topo, mass, temp, velo) in enumerate(draws):
        one = W.Context(g, nlayers=L)
        one.set_topography(topo)
        one.upload_state(0, mass, temp, velo)
        one.fs_set_params(dt, 0.6, 0.8)
        one.fs_upload(W.FS_ETA, etas[m])
        one.set_precision(prec)
        one.set_stage_shape(W.SHAPE_LAYER_PAIRS)
        its.append(one.rk4_split(0, 2, 1e-11, 400))
        sm, st, su = one.download_state(0)
        sl = slice(m * L, (m + 1) * L)
        assert np.array_equal(em[sl][:, own], sm[:, own])
        assert np.array_equal(et[sl][:, own], st[:, own])
        assert np.array_equal(eu[sl][:, :, own_e], su[:, :, own_e])
        assert np.array_equal(ee[m][own], one.fs_download(W.FS_ETA)[own])
        one.close()
    print(f"  iterations: ensemble {it_e}, members {its}", flush=True)
    assert it_e == max(its)
    with pytest.raises(W.WtgError, match="ensemble"):
        ens.elliptic_solve(W.SOLVE_BICGSTAB, 1e-9, 10)
    ens.close()
