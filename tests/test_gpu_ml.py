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
                                                         of their scale
  inverse_wavelet_transform of the reference's wavelets  within 1e-14
Hierarchy J5..J7 (the levels of the J7 ocean configuration), three layers of
the synthetic ocean: one RK4 step bit-exact against the oracle.
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
    assert maxrel(wm[:, ocf], f[f"w0_mass_{l1}"][:, ocf]) < 1e-12
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
