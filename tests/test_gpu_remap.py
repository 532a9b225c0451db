"""Vertical remap of the Lagrangian layers on the GPU (wtg_remap; needs a GPU).

Bars: against the reference's remap_vertical_coordinates (fixtures
trisk_ocean_remap{1,0,2ppm,ppr}_J4: remap_type "1", "2PPM" and "PPR" onto
Chebyshev and "0" onto uniform hybrid coordinates after one RK4 step of the
J4 ocean, 6 layers; PPR also from evenly spaced old layers and with three
layers; "2S", "2W" and "4" onto Chebyshev) within 1e-15 relative; against the oracle on a J7 ocean with 20 layers within 1e-15 (same
evaluation order, no FMA contraction: observed bit for bit); every member of
an ensemble equal to its own single-member run bit for bit; at J8 with 30
layers the column mass and the column buoyancy content are conserved to
rounding (size-independent properties of the finite-volume remap).
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from conftest import golden, maxrel
from helpers import grid_of_fixture, oracle_of

pytestmark = pytest.mark.gpu


def _coords(L, kind):
    k = np.arange(L + 1)
    b = (1.0 + np.cos(k / L * np.pi)) / 2.0 if kind == "chebyshev" else 1.0 - k / L
    return 1.0 - b, b


@pytest.mark.parametrize("case", ["ocean_remap1_J4", "ocean_remap0_J4", "ocean_remap2ppm_J4", "ocean_remapppr_J4",
                                  "ocean_remapppr_uniform_J4", "ocean_remapppr_l3_J4", "ocean_remap2s_J4",
                                  "ocean_remap2w_J4", "ocean_remap4_J4"])
def test_remap_matches_reference(case):
    print(f"remap {case}", flush=True)
    f = golden(f"trisk_{case}.npz")
    g = grid_of_fixture(f)
    ns, L = int(f["nsteps"]), int(f["nlayers"])
    topo = np.nan_to_num(f["topo"])
    m, t, u = (np.nan_to_num(f[f"{v}{ns}"]) for v in ("mass", "temp", "velo"))
    ctx = W.Context(g, nlayers=L, grav=float(f["grav"]), ref_density=float(f["ref_density"]))
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.remap(0, int(str(f["remap_scheme"])), f["a_vert"], f["b_vert"])
    gm, gt, gu = ctx.download_state(0)
    ctx.close()
    oc, oe = g.layout.owned_cells(), g.layout.owned_cells(with_poles=False)
    for k in range(L):
        assert maxrel(gm[k, oc], f["remapped_mass"][k, oc]) <= 1e-15
        assert maxrel(gt[k, oc], f["remapped_temp"][k, oc]) <= 1e-15
        assert maxrel(gu[k, :, oe], f["remapped_velo"][k, :, oe]) <= 1e-15


# the synthetic ocean's layers are uniform (PPR's uniform-spacing branch): every scheme remaps onto Chebyshev
# coordinates
@pytest.mark.parametrize("scheme,kind", [(W.REMAP_PIECEWISE_LINEAR, "chebyshev"), (W.REMAP_PIECEWISE_CONSTANT, "chebyshev"),
                                         (W.REMAP_PARABOLIC_PPM, "chebyshev"), (W.REMAP_PPR, "chebyshev"),
                                         (W.REMAP_PARABOLIC_SPLINE, "chebyshev"), (W.REMAP_PARABOLIC_WENO, "chebyshev"),
                                         (W.REMAP_QUARTIC_WENO, "chebyshev")])
def test_remap_j7_l20_equals_oracle(scheme, kind):
    print(f"remap J7 L20 scheme {scheme} onto {kind} coordinates", flush=True)
    from oracle.oracle import Oracle, geometry_of
    g = W.Grid(level=7, min_level=7, domain_level=2)
    L = 20
    topo, m, t, u = cases.ocean(g, L, seed=5)
    a, b = _coords(L, kind)
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.remap(0, scheme, a, b)
    got = ctx.download_state(0)
    ctx.close()
    o = Oracle(g.layout, geometry_of(g), topo)
    mo, to, uo = m.copy(), t.copy(), u.copy()
    o.ghost_fill(mo, to, uo)
    want = o.remap(mo, to, uo, scheme, a, b)
    oc, oe = g.layout.owned_cells(), g.layout.owned_cells(with_poles=False)
    assert maxrel(got[0][:, oc], want[0][:, oc]) <= 1e-15
    assert maxrel(got[1][:, oc], want[1][:, oc]) <= 1e-15
    assert maxrel(got[2][:, :, oe], want[2][:, :, oe]) <= 1e-15
    assert maxrel(got[0][:, oc], m[:, oc]) > 1e-3  # not a no-op


@pytest.mark.parametrize("scheme", [W.REMAP_PIECEWISE_LINEAR, W.REMAP_PPR])
def test_remap_ensemble_members_equal_single_runs(scheme):
    print(f"remap ensemble of 2, J5 L6, scheme {scheme}", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L = 6
    draws = [cases.ocean(g, L, seed=30 + k) for k in range(2)]
    a, b = _coords(L, "chebyshev")
    ens = W.Context(g, nlayers=L, members=2)
    ens.set_topography(np.stack([d[0] for d in draws]))
    ens.upload_state(0, *(np.concatenate([d[v] for d in draws]) for v in (1, 2, 3)))
    ens.remap(0, scheme, a, b)
    em, et, eu = ens.download_state(0)
    ens.close()
    for k, (topo, m, t, u) in enumerate(draws):
        c = W.Context(g, nlayers=L)
        c.set_topography(topo)
        c.upload_state(0, m, t, u)
        c.remap(0, scheme, a, b)
        sm, st, su = c.download_state(0)
        c.close()
        sl = slice(k * L, (k + 1) * L)
        assert np.array_equal(em[sl], sm) and np.array_equal(et[sl], st) and np.array_equal(eu[sl], su)


@pytest.mark.parametrize("scheme", [W.REMAP_PIECEWISE_LINEAR, W.REMAP_PPR])
def test_remap_j8_l30_conserves_columns(scheme):
    print(f"remap J8 L30 scheme {scheme}: column mass and buoyancy content", flush=True)
    g = W.Grid(level=8, min_level=8, domain_level=2)
    L = 30
    topo, m, t, u = cases.ocean(g, L, seed=8)
    a, b = _coords(L, "chebyshev")
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.remap(0, scheme, a, b)
    gm, gt, gu = ctx.download_state(0)
    ctx.close()
    oc = g.layout.owned_cells()
    assert np.max(np.abs(gm[:, oc].sum(0) - m[:, oc].sum(0)) / m[:, oc].sum(0)) <= 1e-14
    assert np.max(np.abs(gt[:, oc].sum(0) - t[:, oc].sum(0))) / np.max(np.abs(t[:, oc].sum(0))) <= 1e-12
    dz = gm[:, oc] / 1028.0
    depth = dz.sum(0)
    assert np.allclose(dz, np.outer(b[:-1] - b[1:], depth), rtol=1e-12, atol=0)  # the layers are the target's


def test_remap_refusals():
    print("remap refusals", flush=True)
    g = W.Grid(level=4, min_level=4, domain_level=2)
    c = W.Context(g, nlayers=3)
    topo, m, t, u = cases.ocean(g, 3, seed=1)
    c.set_topography(topo)
    c.upload_state(0, m, t, u)
    a, b = _coords(3, "uniform")
    with pytest.raises(W.WtgError):
        c.remap(0, 7, a, b)  # no such remap_type
    with pytest.raises(ValueError):
        c.remap(0, W.REMAP_PIECEWISE_LINEAR, a[:-1], b[:-1])
    c.close()
    # remap4 reads a third layer, remap2S a second one, that the reference leaves unset below that
    for L, scheme in ((2, W.REMAP_QUARTIC_WENO), (1, W.REMAP_PARABOLIC_SPLINE)):
        few = W.Context(g, nlayers=L)
        t2, m2, tt2, u2 = cases.ocean(g, L, seed=1)
        few.set_topography(t2)
        few.upload_state(0, m2, tt2, u2)
        with pytest.raises(W.WtgError):
            few.remap(0, scheme, *_coords(L, "uniform"))
        few.close()
    big = W.Context(g, nlayers=W.REMAP_MAX_LAYERS + 1)
    big.set_topography(topo)
    a, b = _coords(W.REMAP_MAX_LAYERS + 1, "uniform")
    with pytest.raises(W.WtgError):
        big.remap(0, W.REMAP_PIECEWISE_LINEAR, a, b)
    big.close()
