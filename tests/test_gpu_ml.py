"""HIP multi-level explicit step (wtg_ml_*) against the oracle and the
reference (needs a GPU).

Hierarchy J4..J5 of tests/golden/mlstep_J4.npz (the reference's uniform
refinement with zero thresholds):
  trend_ml on both levels, WT_after_step, two RK4 steps   bit-exact against
                                                         the oracle (owned cells,
                                                         pole masters, owned edges)
  the same against the reference                        the bars of
                                                         tests/test_ml_oracle.py
  forward_wavelet_transform                             restricted coarse state
                                                         within 1e-14 of the
                                                         reference's, velocity
                                                         wavelets within 1e-12
                                                         of their scale, scalar
                                                         wavelets within 1e-12
                                                         outside 80 degrees
                                                         (1e-11 with the polar
                                                         caps, where the
                                                         reference's pole copies
                                                         drift)
  inverse_wavelet_transform of the reference's wavelets  within 1e-14
Hierarchy J5..J7 (the levels of the J7 ocean configuration), three layers of
the synthetic ocean: one RK4 step bit-exact against the oracle.

RK4_split on the hierarchy with the FMG free surface (wtg_ml_rk4_split):
  J4..J5 of tests/golden/mlsplit_J4.npz, each of two    the oracle's V-cycle count;
  steps from the reference's state before it             state within 1e-12 and free
                                                         surface within 1e-11 of the
                                                         oracle's (the GPU's norms reduce
                                                         in another order, so FMG's
                                                         iterates differ in the last bits);
                                                         the reference within the bars of
                                                         tests/test_mlsplit_oracle.py
                                                         (x 10)
  J5..J7, 20 layers of the synthetic ocean, FMG to      the oracle's V-cycle count; state
  1e-9 (the J7 ocean configuration)                      within 1e-11 and free surface
                                                         within 1e-10 of the oracle's
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden, maxrel
from oracle.oracle import MlOracle
from test_ml_oracle import _stage_state, outside_caps, owned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    f = golden("mlstep_J4.npz")
    ls, le = int(f["level_start"]), int(f["level_end"])
    levels = list(range(ls, le + 1))
    grids = [W.Grid(level=l, min_level=int(f["min_level"]), domain_level=int(f["domain_level"]),
                    radius=float(f["radius"]), omega=float(f["omega"])) for l in levels]
    topos = [np.nan_to_num(f[f"topo_{l}"]) for l in levels]
    o = MlOracle(grids, topos, grav=float(f["grav"]), rho0=float(f["ref_density"]))
    s0 = [[np.nan_to_num(f[f"s0_{v}_{l}"]) for l in levels] for v in ("mass", "temp", "velo")]
    o.ghost_fill(*s0)
    L = int(f["nlayers"])
    ctxs = [W.Context(g, nlayers=L, grav=float(f["grav"]), ref_density=float(f["ref_density"]), nslots=2)
            for g in grids]
    for c, t in zip(ctxs, topos):
        c.set_topography(t)
    ml = W.MultiLevel(ctxs)
    yield f, levels, grids, o, s0, ctxs, ml
    ml.close()


def _same(got, want, grids):
    """bit-exact on owned cells, pole masters and owned edges of every level"""
    for i, g in enumerate(grids):
        oc, oe = owned(g)
        for a, b in zip(got[:2], want[:2]):
            assert np.array_equal(a[i][:, oc], b[i][:, oc]), f"level {i}: scalar differs"
        assert np.array_equal(got[2][i][:, :, oe], want[2][i][:, :, oe]), f"level {i}: velocity differs"


def test_trend_ml_equals_oracle(setup):
    f, levels, grids, o, s0, ctxs, ml = setup
    print("trend_ml J4..J5 hierarchy", flush=True)
    ml.upload(0, *s0)
    ml.trend(0, 1)
    got = ml.download(1)
    want = o.trend(*s0)
    _same(got, want, grids)
    for i, (l, g) in enumerate(zip(levels, grids)):
        oc, oe = owned(g)
        assert maxrel(got[0][i][:, oc], f[f"t0_mass_{l}"][:, oc]) < 1e-13


def test_wt_after_step_equals_oracle(setup):
    f, levels, grids, o, s0, ctxs, ml = setup
    print("WT_after_step J4..J5 hierarchy", flush=True)
    st = _stage_state(f, levels, grids, o, s0)
    ml.upload(0, *st)
    ml.wt_after_step(0)
    _same(ml.download(0), o.wt_after_step(*st), grids)


def test_rk4_two_steps_equals_oracle_and_reference(setup):
    f, levels, grids, o, s0, ctxs, ml = setup
    print("RK4 on the J4..J5 hierarchy, two steps", flush=True)
    ml.upload(0, *s0)
    want = s0
    for s in (1, 2):
        ml.rk4(0, float(f["dt"]), 1)
        got = ml.download(0)
        want = o.rk4(*want, float(f["dt"]), 1)
        _same(got, want, grids)
        for i, (l, g) in enumerate(zip(levels, grids)):
            oc, oe = outside_caps(g, 75.0)
            ref_m, ref_u = f[f"s{s:02d}_mass_{l}"], f[f"s{s:02d}_velo_{l}"]
            assert np.abs(got[0][i][:, oc] - ref_m[:, oc]).max() / np.abs(ref_m[:, owned(g)[0]]).max() < 1e-12
            assert np.abs(got[2][i][:, :, oe] - ref_u[:, :, oe]).max() / np.abs(ref_u[:, :, owned(g)[1]]).max() < 1e-12


def test_forward_and_inverse_vs_reference(setup):
    f, levels, grids, o, s0, ctxs, ml = setup
    print("forward / inverse wavelet transform J4..J5", flush=True)
    l0, l1 = levels
    # forward: the finest state, a coarse level to be overwritten
    junk = [np.zeros_like(a) for a in (s0[0][0], s0[1][0], s0[2][0])]
    ml.upload(0, [junk[0], s0[0][1]], [junk[1], s0[1][1]], [junk[2], s0[2][1]])
    ml.forward(0)
    m, t, u = ml.download(0)
    oc, oe = owned(grids[0])
    keep = oc[oc != grids[0].layout.north_pole]  # the reference's north pole copy drifts
    assert maxrel(m[0][:, keep], f[f"s0_mass_{l0}"][:, keep]) < 1e-14
    assert maxrel(u[0][:, :, oe], f[f"s0_velo_{l0}"][:, :, oe]) < 1e-14
    wm, wt, wu = ml.wavelets(1)
    ocf, oef = owned(grids[1])
    assert maxrel(wu[:, :, oef], f[f"w0_velo_{l1}"][:, :, oef]) < 1e-12
    # scalar wavelets within 5 degrees of a pole read the reference's drifting pole copies
    cap, _ = outside_caps(grids[1], 80.0)
    assert maxrel(wm[:, cap], f[f"w0_mass_{l1}"][:, cap]) < 1e-12
    assert maxrel(wm[:, ocf], f[f"w0_mass_{l1}"][:, ocf]) < 1e-11
    # inverse of the reference's wavelets from the reference's coarse state
    ml.upload(0, *s0)
    ml.set_wavelets(1, np.nan_to_num(f[f"w0_mass_{l1}"]), wt, np.nan_to_num(f[f"w0_velo_{l1}"]))
    ml.inverse(0)
    _, _, u = ml.download(0)
    assert maxrel(u[1][:, :, oef], f[f"i0_velo_{l1}"][:, :, oef]) < 1e-14


def test_rk4_j5_j7_ocean_equals_oracle():
    from wavetrisk_hydrostatic_amd import cases
    print("RK4 on the J5..J7 hierarchy, 3 layers, one step", flush=True)
    L = 3
    grids = [W.Grid(level=l, min_level=5, domain_level=2) for l in (5, 6, 7)]
    states, topos = [], []
    for g in grids:
        topo, m, t, u = cases.ocean(g, L, seed=11)
        topos.append(topo)
        states.append((m, t, u))
    o = MlOracle(grids, topos)
    s0 = [[s[0] for s in states], [s[1] for s in states], [s[2] for s in states]]
    o.ghost_fill(*s0)
    ctxs = [W.Context(g, nlayers=L, nslots=2) for g in grids]
    for c, t in zip(ctxs, topos):
        c.set_topography(t)
    ml = W.MultiLevel(ctxs)
    try:
        ml.upload(0, *s0)
        ml.forward(0)  # coarse levels from the finest, as the reference initialises them
        s1 = ml.download(0)
        o.ghost_fill(*s1)
        ml.rk4(0, 300.0, 1)
        _same(ml.download(0), o.rk4(*s1, 300.0, 1), grids)
    finally:
        ml.close()


def _split_contexts(grids, topos, L, grav, rho0, dt, th1, th2):
    ctxs = []
    for g, t in zip(grids, topos):
        c = W.Context(g, nlayers=L, grav=grav, ref_density=rho0, nslots=2)
        c.set_topography(t)
        c.fs_set_params(dt, th1, th2)
        ctxs.append(c)
    return ctxs


def test_rk4_split_hierarchy_vs_oracle_and_reference():
    from test_mlsplit_oracle import errors, load, mg_of, state_before
    print("RK4_split on the J4..J5 hierarchy, FMG free surface, two steps", flush=True)
    f, levels, grids, topos = load()
    grav, rho0 = float(f["grav"]), float(f["ref_density"])
    o = MlOracle(grids, topos, grav=grav, rho0=rho0)
    ctxs = _split_contexts(grids, topos, int(f["nlayers"]), grav, rho0, float(f["dt"]), float(f["theta1"]),
                           float(f["theta2"]))
    h = W.Hierarchy(ctxs)
    ml = W.MultiLevel(ctxs)
    tol = float(f["tol"])
    p = W.MgParams.defaults(coarse_tol=tol, fine_tol=tol, coarse_iter=400, max_vcycle=50)
    try:
        for step in (1, 2):
            st, eta = state_before(f, levels, step)
            o.ghost_fill(*st)
            ml.upload(0, *st)
            for c, e in zip(ctxs, eta):
                c.fs_upload(W.FS_ETA, e)
            it = ml.rk4_split(h, 0, 1, p)
            m, t, u = ml.download(0)
            e = [c.fs_download(W.FS_ETA) for c in ctxs]
            mo, to, uo, eo, ito = o.rk4_split(mg_of(f, grids, topos), *st, eta)
            print(f"step {step}: FMG V-cycles GPU {it}, oracle {list(ito)}", flush=True)
            assert it == int(ito.sum())
            for i, g in enumerate(grids):
                oc, oe = owned(g)
                assert maxrel(m[i][:, oc], mo[i][:, oc]) < 1e-12
                assert maxrel(t[i][:, oc], to[i][:, oc]) < 1e-12
                assert maxrel(u[i][:, :, oe], uo[i][:, :, oe]) < 1e-12
                assert maxrel(e[i][oc], eo[i][oc]) < 1e-11
            for l, (em, eu, ee) in zip(levels, errors(f, levels, grids, step, m, u, e)):
                assert em < 1e-12 and eu < 1e-12 and ee < 1e-11, (l, em, eu, ee)
    finally:
        ml.close()
        h.close()


def test_rk4_split_j5_j7_l20_ocean_vs_oracle():
    """the J7 ocean configuration's hierarchy (J5..J7, 20 layers), FMG to 1e-9"""
    from wavetrisk_hydrostatic_amd import cases
    from oracle.oracle import MgOracle
    print("RK4_split on the J5..J7 hierarchy, 20 layers, FMG to 1e-9, one step", flush=True)
    L, dt, th1, th2 = 20, 300.0, 0.6, 0.8
    grids = [W.Grid(level=l, min_level=5, domain_level=2) for l in (5, 6, 7)]
    states, topos = [], []
    for g in grids:
        topo, m, t, u = cases.ocean(g, L, seed=11)
        topos.append(topo)
        states.append((m, t, u))
    ctxs = _split_contexts(grids, topos, L, 9.80616, 1028.0, dt, th1, th2)
    h = W.Hierarchy(ctxs)
    ml = W.MultiLevel(ctxs)
    try:
        ml.upload(0, [s[0] for s in states], [s[1] for s in states], [s[2] for s in states])
        ml.forward(0)  # coarse levels from the finest
        s1 = list(ml.download(0))
        o = MlOracle(grids, topos)
        o.ghost_fill(*s1)
        eta = [s1[0][i].sum(0) / 1028.0 + topos[i] for i in range(3)]
        for c, e in zip(ctxs, eta):
            c.fs_upload(W.FS_ETA, e)
        p = W.MgParams.defaults(coarse_tol=1e-9, fine_tol=1e-9, coarse_iter=400, max_vcycle=50)
        it = ml.rk4_split(h, 0, 1, p)
        m, t, u = ml.download(0)
        e = [c.fs_download(W.FS_ETA) for c in ctxs]
        mg = MgOracle(grids, topos, dt=dt, theta1=th1, theta2=th2, coarse_tol=1e-9, fine_tol=1e-9, coarse_iter=400,
                      max_vcycle=50)
        mo, to, uo, eo, ito = o.rk4_split(mg, *s1, eta)
        print(f"FMG V-cycles GPU {it}, oracle {list(ito)}", flush=True)
        assert it == int(ito.sum())
        for i, g in enumerate(grids):
            oc, oe = owned(g)
            assert maxrel(m[i][:, oc], mo[i][:, oc]) < 1e-11
            assert maxrel(u[i][:, :, oe], uo[i][:, :, oe]) < 1e-11
            assert maxrel(e[i][oc], eo[i][oc]) < 1e-10
            assert np.isfinite(m[i][:, oc]).all() and np.isfinite(u[i][:, :, oe]).all()
    finally:
        ml.close()
        h.close()


def test_rk3_split_hierarchy_vs_oracle():
    """RK3_split on the J4..J5 hierarchy (wtg_ml_rk_split, 3 stages) from
    the reference's initial state: the oracle's V-cycle count, state within
    1e-12 and free surface within 1e-11 of the oracle's"""
    from test_mlsplit_oracle import load, mg_of, state_before
    print("RK3_split on the J4..J5 hierarchy", flush=True)
    f, levels, grids, topos = load()
    grav, rho0 = float(f["grav"]), float(f["ref_density"])
    o = MlOracle(grids, topos, grav=grav, rho0=rho0)
    ctxs = _split_contexts(grids, topos, int(f["nlayers"]), grav, rho0, float(f["dt"]), float(f["theta1"]),
                           float(f["theta2"]))
    h = W.Hierarchy(ctxs)
    ml = W.MultiLevel(ctxs)
    tol = float(f["tol"])
    p = W.MgParams.defaults(coarse_tol=tol, fine_tol=tol, coarse_iter=400, max_vcycle=50)
    try:
        st, eta = state_before(f, levels, 1)
        o.ghost_fill(*st)
        ml.upload(0, *st)
        for c, e in zip(ctxs, eta):
            c.fs_upload(W.FS_ETA, e)
        it = ml.rk4_split(h, 0, 1, p, nstages=3)
        m, t, u = ml.download(0)
        e = [c.fs_download(W.FS_ETA) for c in ctxs]
        mo, to, uo, eo, ito = o.rk4_split(mg_of(f, grids, topos), *st, eta, nstages=3)
        assert it == int(ito.sum())
        for i, g in enumerate(grids):
            oc, oe = owned(g)
            assert maxrel(m[i][:, oc], mo[i][:, oc]) < 1e-12
            assert maxrel(u[i][:, :, oe], uo[i][:, :, oe]) < 1e-12
            assert maxrel(e[i][oc], eo[i][oc]) < 1e-11
    finally:
        ml.close()
        h.close()
