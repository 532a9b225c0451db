"""The column physics at the largest column the kernels hold
(WTG_REMAP_MAX_LAYERS = 64 layers; needs a GPU): every remap scheme and both
vertical-diffusion closures on a J4 ocean with 64 layers against the oracle,
bit for bit (the analytic eddy profiles within 1e-14: exp).  The private
column arrays are filled to their last entry here, so an index one past a
column would show as a difference."""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from conftest import maxrel

pytestmark = pytest.mark.gpu

L = W.REMAP_MAX_LAYERS


@pytest.fixture(scope="module")
def setup():
    from oracle.oracle import Oracle, geometry_of
    g = W.Grid(level=4, min_level=4, domain_level=2)
    topo, m, t, u = cases.ocean(g, L, seed=64)
    o = Oracle(g.layout, geometry_of(g), topo)
    return g, topo, m, t, u, o


@pytest.mark.parametrize("scheme", range(7))
def test_remap_64_layers_equals_oracle(setup, scheme):
    print(f"remap J4 L64 scheme {scheme}", flush=True)
    g, topo, m, t, u, o = setup
    k = np.arange(L + 1)
    b = (1.0 + np.cos(k / L * np.pi)) / 2.0
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.remap(0, scheme, 1.0 - b, b)
    got = ctx.download_state(0)
    ctx.close()
    st = (m.copy(), t.copy(), u.copy())
    o.ghost_fill(*st)
    want = o.remap(*st, scheme, 1.0 - b, b)
    oc, oe = g.layout.owned_cells(), g.layout.owned_cells(with_poles=False)
    assert np.array_equal(got[0][:, oc], want[0][:, oc]) and np.array_equal(got[1][:, oc], want[1][:, oc])
    assert np.array_equal(got[2][:, :, oe], want[2][:, :, oe])


@pytest.mark.parametrize("tke", [0, 1])
def test_vertical_diffusion_64_layers_equals_oracle(setup, tke):
    print(f"vertical diffusion J4 L64 tke_closure={tke}", flush=True)
    g, topo, m, t, u, o = setup
    NC = g.layout.ncells
    rng = np.random.default_rng(64)
    forcing = dict(bot=rng.uniform(-3e-7, 3e-7, NC), top=rng.uniform(-1e-6, 1e-6, NC), tau=rng.uniform(0.02, 0.2, NC),
                   wind=rng.uniform(-2e-4, 2e-4, (3, NC)))
    coef = (np.full((L + 1, NC), 1.2e-5), np.full((L + 1, NC), 1.2e-4), rng.uniform(1e-6, 3e-5, (L, NC)))
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.set_vdiff(tke_closure=tke, bottom_friction=4e-4)
    ctx.vdiff_set_coefficients(*coef)
    ctx.vdiff_set_forcing(forcing["bot"], forcing["top"], forcing["tau"], forcing["wind"])
    for c in (1, 2):
        ctx.vertical_diffusion(0, 1200.0, c)
    got = ctx.download_state(0) + ctx.vdiff_coefficients()
    ctx.close()
    st = (m.copy(), t.copy(), u.copy())
    o.ghost_fill(*st)
    for c in (1, 2):
        res = o.vertical_diffusion(*st, 1200.0, c, forcing, *coef, tke_closure=tke, bottom_friction=4e-4)
        st, coef = res[:3], res[3:]
    oc, oe = g.layout.owned_cells(), g.layout.owned_cells(with_poles=False)
    err = [maxrel(got[1][:, oc], res[1][:, oc]), maxrel(got[2][:, :, oe], res[2][:, :, oe])] + \
          [maxrel(got[k][:, oc], res[k][:, oc]) for k in (3, 4, 5)]
    assert max(err) <= (0.0 if tke else 1e-14), err
